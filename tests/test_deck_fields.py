"""The generated fields of the benchmark decks are reproducible: their SHA-256 equal the
committed ones (tests/golden/field_checksums.json, written by make_field_checksums.py), and
the recipes' stated ranges hold (config 4: nx, ny in 32..96, nz in 4..8, sigma in
[0.5, 2], mobility ratio in [0.5, 50], Corey exponents in [1.5, 4])."""
import json
import os

import numpy as np

from ressim_amd import decks

HERE = os.path.dirname(os.path.abspath(__file__))


def test_field_checksums_match_the_committed_ones():
    import importlib.util

    spec = importlib.util.spec_from_file_location("mk", os.path.join(HERE, "golden", "make_field_checksums.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    want = json.load(open(os.path.join(HERE, "golden", "field_checksums.json")))
    got = mk.checksums()
    assert sorted(got) == sorted(want)
    bad = [k for k in want if got[k] != want[k]]
    assert not bad, bad[:5]


def test_ensemble_recipe_ranges():
    for m in range(256):
        d = decks.two_phase_member(m, n_steps=1)
        assert 32 <= d.nx <= 96 and 32 <= d.ny <= 96 and 4 <= d.nz <= 8
        assert 0.5 <= d.info["sigma"] <= 2.0 and 0.5 <= d.info["mobility_ratio"] <= 50.0
        assert 1.5 <= d.info["n_w"] <= 4.0 and 1.5 <= d.info["n_o"] <= 4.0


def test_grf_statistics():
    f = decks.gaussian_random_field((16, 64, 64), 8.0, 1.5, seed=1)
    assert abs(f.mean()) < 1e-12 and abs(f.std() - 1.5) < 1e-12
    # exponential covariance: the lag-1 correlation along x is near exp(-1/8)
    c = np.mean(f[:, :, :-1] * f[:, :, 1:]) / f.var()
    assert 0.75 < c < 0.95
