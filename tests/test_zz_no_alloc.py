"""rsf_cnv_mb_noalloc, the entry point that allocates nothing, against rsf_cnv_mb (the same
kernels with an allocated workspace) through the C ABI: bit-identical on two calls in a row,
the second with a smaller input, so the second call reads a workspace that still holds the
first call's records; the size limit is enforced."""
import ctypes
import sys

import numpy as np
import pytest

from ressim_amd import simulator as rsim

pytestmark = pytest.mark.gpu


def _case(n, seed):
    rng = np.random.default_rng(seed)
    pv = rng.uniform(10.0, 200.0, size=n)
    fvf = np.stack([rng.uniform(0.98, 1.02, n), rng.uniform(1.1, 1.3, n), rng.uniform(0.004, 0.01, n)], axis=1)
    R = rng.standard_normal((n, 3)) * 1e-3 * pv[:, None]
    hot = rng.random(n) < 0.02
    R[hot, 1] = 0.5 * pv[hot]
    return R.ravel(), pv, fvf.ravel()


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def test_noalloc_matches_allocating_entry_point():
    import torch

    for n, seed, relax in ((400_000, 1, False), (3_000, 2, True), (1, 3, False)):
        R, pv, fvf = (_dev(a) for a in _case(n, seed))
        a = rsim.cnv_mb(R, pv, fvf, relax_final=relax)
        b = rsim.cnv_mb_noalloc(R, pv, fvf, relax_final=relax)
        print(f"[noalloc] n={n} cnv={b['cnv']} mb={b['mb']} viol={b['violating_pv_fraction']:.4f}", flush=True)
        np.testing.assert_array_equal(b["cnv"], a["cnv"])
        np.testing.assert_array_equal(b["mb"], a["mb"])
        for k in ("violating_pv_fraction", "pv_rule_relaxes", "would_accept", "would_accept_strict"):
            assert b[k] == a[k], k
    # on a stream of the caller's, with the inputs written on that stream
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        R, pv, fvf = (_dev(a) for a in _case(1234, 4))
        b = rsim.cnv_mb_noalloc(R, pv, fvf, stream=s)
    a = rsim.cnv_mb(R, pv, fvf)
    np.testing.assert_array_equal(b["cnv"], a["cnv"])
    np.testing.assert_array_equal(b["mb"], a["mb"])


def test_noalloc_rejects_sizes_beyond_its_maximum():
    import torch

    out = torch.empty(10, dtype=torch.float64, device="cuda")
    x = torch.zeros(3, dtype=torch.float64, device="cuda")
    rc = rsim.lib().rsf_cnv_mb_noalloc(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                       ctypes.c_void_p(x.data_ptr()), (1 << 28) + 1, 0,
                                       ctypes.c_void_p(out.data_ptr()), None)
    assert rc == rsim.RSF_ERR_INVALID
    buf = ctypes.create_string_buffer(256)
    rsim.lib().rsf_cnv_mb_last_error(buf, 256)
    assert b"outside" in buf.value, buf.value
    sys.stdout.flush()
