"""Batched sensitivity ensembles: many independent FIM models stepped concurrently
on one GPU.

Every member is its own simulator handle with its own HIP stream, so kernels of
different members overlap on the device; host threads drive the members (the C ABI
releases the interpreter lock during each call). All members are built and their
device state is set up before the timed region; the timed region covers every
report step of every member.
"""
from __future__ import annotations

import concurrent.futures as cf
import threading
import time

from .simulator import ReservoirSimulator


def build_members(member_decks, solver_options=None):
    """Simulators for the decks, with device buffers and initial state resident."""
    sims = []
    for d in member_decks:
        s = d.build(ReservoirSimulator)
        if solver_options:
            s.set_solver_options(**solver_options)
        s.debug_initial_state()  # allocates the device state and uploads the model
        sims.append(s)
    return sims


def run_ensemble(member_decks, sims, n_threads=8):
    """Step every member through its schedule; members are taken from a shared queue
    by `n_threads` host threads. Returns per-member status and the wall time."""
    order = sorted(range(len(sims)), key=lambda k: -member_decks[k].n_cells)  # largest first
    lock = threading.Lock()
    nxt = [0]
    status = [None] * len(sims)

    def worker():
        while True:
            with lock:
                if nxt[0] >= len(order):
                    return
                k = order[nxt[0]]
                nxt[0] += 1
            d, s = member_decks[k], sims[k]
            rc = 0
            for dt in d.schedule:
                rc = s.step(dt)
                if rc:
                    break
            status[k] = rc

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=n_threads) as ex:
        for f in [ex.submit(worker) for _ in range(n_threads)]:
            f.result()
    elapsed = time.perf_counter() - t0
    cell_steps = sum(d.n_cells * len(d.schedule) for d, st in zip(member_decks, status) if st == 0)
    return dict(elapsed=elapsed, status=status, cell_steps=cell_steps,
                cell_steps_per_s=cell_steps / elapsed if elapsed > 0 else 0.0)
