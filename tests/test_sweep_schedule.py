"""The schedule and index arithmetic of the one-launch natural-order sweeps and
factorizations (ressim_amd/csrc/rsf_sweep.hpp: sweep_tiles, k_ilu_sweep,
k_ilu_factor_sweep), restated over every (tile, step, thread) with numpy.

These kernels wait on one another's published halo entries, so their correctness rests on
invariants that a run on a few decks does not exercise for every tile shape:
  * the march of all tiles visits every cell of the grid exactly once;
  * every load the kernels issue (prefetches of clamped steps, halo entries of threads
    that do not use them) stays inside its buffer, and the natural index nat0 + row of a
    cell inside the grid is that cell;
  * a thread that takes a neighbour from a halo entry finds that the entry is written by
    the tile before it, at the step and with the tag it expects, for the neighbouring cell;
    the tile it waits on has a lower workgroup index (dispatched earlier), so no wait can
    depend on a workgroup that is not running;
  * a thread that takes a neighbour from LDS reads the thread that handled the
    neighbouring cell one step earlier.
The index formulas below are those of the kernels; a change there must be made here too.
"""
import numpy as np
import pytest


def sweep_tiles(nx, ny, nz, lanes):
    """rsf_sweep.hpp sweep_tiles: ty x tz lines per workgroup, grown by doubling."""
    ty = tz = 1
    while ty * tz < lanes:
        cy, cz = -(-ny // ty), -(-nz // tz)
        if ty < ny and (cy >= cz or tz >= nz):
            ty *= 2
        elif tz < nz:
            tz *= 2
        else:
            break
    return dict(ty=ty, tz=tz, nty=-(-ny // ty), ntz=-(-nz // tz), steps=nx + ty + tz - 2)


def schedule(nx, ny, nz, st, bwd):
    """Per (block, step, thread): tile, lane coordinates, cell and flags, as the kernels."""
    ty, tz, nty, ntz, S = st["ty"], st["tz"], st["nty"], st["ntz"], st["steps"]
    nT, NL = nty * ntz, ty * tz
    blk = np.arange(nT)[:, None, None]
    t = np.arange(S)[None, :, None]
    L = np.arange(NL)[None, None, :]
    T = nT - 1 - blk if bwd else blk
    tY, tZ = T % nty, T // nty
    a, b = L % ty, L // ty
    y, z = tY * ty + a, tZ * tz + b
    line = (y < ny) & (z < nz)
    row = S - 1 - t if bwd else t
    x = row - a - b
    valid = line & (x >= 0) & (x < nx)
    nat0 = nx * (y + ny * z) - a - b
    return dict(T=T, tY=tY, tZ=tZ, t=t, a=a, b=b, x=x, y=y, z=z, row=row, valid=valid, nat0=nat0, nT=nT, NL=NL,
                S=S)


SHAPES = [(12, 10, 6), (9, 7, 5), (11, 9, 4), (7, 1, 1), (1, 5, 3), (40, 33, 17), (64, 64, 16), (3, 130, 2)]
LANES = [1, 2, 8, 16, 32, 64, 256]


@pytest.mark.parametrize("bwd", [False, True], ids=["forward", "backward"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_march_covers_grid_and_halo_waits_are_published_earlier(shape, bwd):
    nx, ny, nz = shape
    n = nx * ny * nz
    for lanes in LANES:
        st = sweep_tiles(nx, ny, nz, lanes)
        ty, tz, nty = st["ty"], st["tz"], st["nty"]
        s = schedule(nx, ny, nz, st, bwd)
        T, t, a, b, valid = s["T"], s["t"], s["a"], s["b"], s["valid"]
        nT, S, NL = s["nT"], s["S"], s["NL"]
        V = np.broadcast_to(valid, (nT, S, NL))
        # every cell once, and nat0 + row is the cell's natural index
        nat = np.broadcast_to(s["nat0"] + s["row"], V.shape)[V]
        cell = np.broadcast_to(s["x"] + nx * (s["y"] + ny * s["z"]), V.shape)[V]
        assert np.array_equal(nat, cell)
        assert np.array_equal(np.sort(nat), np.arange(n)), (shape, lanes)
        # producers: (tile, step, lane b or a) -> written entry; tags are step + 1
        hy_line = (s["y"] + 1 < ny) if bwd else (s["y"] > 0)
        hz_line = (s["z"] + 1 < nz) if bwd else (s["z"] > 0)
        out_y = ((a == 0) & (s["tY"] > 0)) if bwd else ((a == ty - 1) & (s["tY"] + 1 < nty))
        out_z = ((b == 0) & (s["tZ"] > 0)) if bwd else ((b == tz - 1) & (s["tZ"] + 1 < st["ntz"]))
        in_y = hy_line & ((a == ty - 1) if bwd else (a == 0))
        in_z = hz_line & ((b == tz - 1) if bwd else (b == 0))
        blk_of_T = (nT - 1 - np.arange(nT)) if bwd else np.arange(nT)  # workgroup index of tile T
        for d, (out_e, in_e, tw, lane_w, plane) in enumerate(
                [(out_y, in_y, ty, b, nT * S * tz), (out_z, in_z, tz, a, nT * S * ty)]):
            stride = tz if d == 0 else ty
            Tu = (T + (1 if d == 0 else nty)) if bwd else (T - (1 if d == 0 else nty))
            prod = np.broadcast_to(valid & out_e, V.shape)
            pidx = np.broadcast_to((T * S + t) * stride + lane_w, V.shape)[prod]
            ptag = np.broadcast_to(t + 1, V.shape)[prod]
            pcell = np.broadcast_to(s["x"] + nx * (s["y"] + ny * s["z"]), V.shape)[prod]
            pblk = np.broadcast_to(blk_of_T[T], V.shape)[prod]
            assert pidx.min(initial=0) >= 0 and pidx.max(initial=-1) < plane
            assert len(np.unique(pidx)) == len(pidx)  # each entry written once per launch
            written = dict(zip(pidx.tolist(), zip(ptag.tolist(), pcell.tolist(), pblk.tolist())))
            cons = np.broadcast_to(valid & in_e, V.shape)
            cidx = np.broadcast_to((Tu * S + t + tw - 1) * stride + lane_w, V.shape)[cons]
            ctag = np.broadcast_to(t + tw, V.shape)[cons]
            step = (nx * ny) if d == 1 else nx
            ncell = np.broadcast_to(s["x"] + nx * (s["y"] + ny * s["z"]), V.shape)[cons] + (step if bwd else -step)
            cblk = np.broadcast_to(blk_of_T[T], V.shape)[cons]
            # the prefetch reads the clamped row: for a thread that uses the entry it is the same
            ridx = np.clip(np.broadcast_to(Tu * S + t + tw - 1, V.shape)[cons], 0, nT * S - 1) * stride + \
                np.broadcast_to(lane_w, V.shape)[cons]
            assert np.array_equal(ridx, cidx)
            for ci, tg, nc_, cb in zip(cidx.tolist(), ctag.tolist(), ncell.tolist(), cblk.tolist()):
                assert ci in written, (shape, lanes, d)
                wtag, wcell, wblk = written[ci]
                assert wtag == tg and wcell == nc_ and wblk < cb, (shape, lanes, d)


@pytest.mark.parametrize("bwd", [False, True], ids=["forward", "backward"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_lds_neighbours_and_load_bounds(shape, bwd):
    nx, ny, nz = shape
    n = nx * ny * nz
    for lanes in LANES:
        st = sweep_tiles(nx, ny, nz, lanes)
        ty, tz = st["ty"], st["tz"]
        s = schedule(nx, ny, nz, st, bwd)
        nT, S, NL = s["nT"], s["S"], s["NL"]
        L = np.arange(NL)
        lane_y = np.clip(L + 1 if bwd else L - 1, 0, NL - 1)
        lane_z = np.clip(L + ty if bwd else L - ty, 0, NL - 1)
        assert lane_y.min() >= 0 and lane_y.max() < NL and lane_z.max() < NL
        a, b = L % ty, L // ty
        cellx = np.broadcast_to(s["x"], (nT, S, NL))
        celly = np.broadcast_to(s["y"], (nT, S, NL))
        cellz = np.broadcast_to(s["z"], (nT, S, NL))
        V = np.broadcast_to(s["valid"], (nT, S, NL))
        hy_line = (celly + 1 < ny) if bwd else (celly > 0)
        hz_line = (cellz + 1 < nz) if bwd else (cellz > 0)
        edge_y = (a == ty - 1) if bwd else (a == 0)
        edge_z = (b == tz - 1) if bwd else (b == 0)
        for lane_n, line_ok, edge, dy, dz in ((lane_y, hy_line, edge_y, 1, 0), (lane_z, hz_line, edge_z, 0, 1)):
            use = V & line_ok & ~edge[None, None, :]
            blk, tt, ll = np.nonzero(use)
            assert (tt >= 1).all()  # the neighbour was handled one step earlier, in this block
            sgn = 1 if bwd else -1
            assert np.array_equal(cellx[blk, tt - 1, lane_n[ll]], cellx[blk, tt, ll])
            assert np.array_equal(celly[blk, tt - 1, lane_n[ll]], celly[blk, tt, ll] + sgn * dy)
            assert np.array_equal(cellz[blk, tt - 1, lane_n[ll]], cellz[blk, tt, ll] + sgn * dz)
            assert V[blk, tt - 1, lane_n[ll]].all()
        # prefetches of steps up to the padded end + the prefetch depth read clamped rows
        for depth in (1, 3, 4, 8):
            s_pad = -(-S // depth) * depth
            tq = np.minimum(np.arange(s_pad + depth), S - 1)
            rows = (S - 1 - tq) if bwd else tq
            assert rows.min() >= 0 and rows.max() <= S - 1
            entry = (np.arange(nT)[:, None, None] * S + rows[None, :, None]) * NL + L[None, None, :]
            assert entry.min() >= 0 and entry.max() < nT * S * NL
        # the factorization's natural-order operand loads (clamped into the grid)
        nat = np.clip(np.broadcast_to(s["nat0"] + s["row"], (nT, S, NL)), 0, n - 1)
        for off in (1, nx, nx * ny):
            assert np.maximum(nat - off, 0).min() >= 0
        assert nat.max() < n
