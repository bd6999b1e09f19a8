"""Writes field_checksums.json: the SHA-256 of every generated FP64 field of the benchmark
decks (ln k and porosity of configs 2, 3 and 5 and of the 256 ensemble members), so that a
change in a generator (seed handling, FFT padding, normalisation) shows up as a failed test.

    python tests/golden/make_field_checksums.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from ressim_amd import decks  # noqa: E402


def checksums():
    out = {}
    for d in (decks.black_oil_65k(1), decks.black_oil_1m(1), decks.black_oil_adjoint(1)):
        out[d.name] = {"lnk_sha256": d.info["lnk_sha256"], "phi_sha256": d.info["phi_sha256"]}
    for m in range(256):
        d = decks.two_phase_member(m, n_steps=1)
        out[d.name] = {"lnk_sha256": d.info["lnk_sha256"], "phi_sha256": d.info["phi_sha256"],
                       "dims": [d.nx, d.ny, d.nz]}
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "field_checksums.json"), "w") as f:
        json.dump(checksums(), f, indent=1, sort_keys=True)
