"""Synthetic decks for the FIM timestep workloads.

A deck is a grid plus an ordered list of setter calls that use the
reference simulator's method names (rs/frontend.rs:36-1325), so the same
deck drives the HIP simulator (``ressim_amd.ReservoirSimulator``) and any
other object exposing those names. Every generator is seeded; the field
recipes (lognormal Gaussian random fields, correlation PVT, Leverett-J
scaling) are this package's own, documented with each builder.
"""
from __future__ import annotations

import dataclasses
import hashlib
from typing import Any, Callable, List, Optional, Tuple

import numpy as np

DARCY = 8.5269888e-3


@dataclasses.dataclass
class Deck:
    name: str
    nx: int
    ny: int
    nz: int
    porosity0: float
    calls: List[Tuple[str, tuple]]
    schedule: List[float]
    info: dict = dataclasses.field(default_factory=dict)

    @property
    def n_cells(self) -> int:
        return self.nx * self.ny * self.nz

    def apply(self, sim) -> None:
        for name, args in self.calls:
            getattr(sim, name)(*args)

    def build(self, factory: Callable[..., Any]):
        sim = factory(self.nx, self.ny, self.nz, self.porosity0)
        self.apply(sim)
        return sim


def sha256_f64(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


# --------------------------------------------------------------------------- fields
def gaussian_random_field(shape, corr_len_cells: float, sigma: float, seed: int) -> np.ndarray:
    """Stationary Gaussian field with exponential covariance sigma^2 exp(-r/L).

    Circulant embedding on a periodic grid padded to twice each extent: the
    covariance is sampled on the torus, its FFT gives the (clipped
    non-negative) eigenvalues, and white noise is coloured by their square
    root. The returned array is indexed [k, j, i] (natural order k*nx*ny +
    j*nx + i after ravel).
    """
    nz, ny, nx = shape
    pz, py, px = (max(2 * s, 2) for s in shape)
    rng = np.random.default_rng(seed)
    z = np.minimum(np.arange(pz), pz - np.arange(pz))[:, None, None]
    y = np.minimum(np.arange(py), py - np.arange(py))[None, :, None]
    x = np.minimum(np.arange(px), px - np.arange(px))[None, None, :]
    r = np.sqrt(x * x + y * y + z * z)
    cov = np.exp(-r / corr_len_cells)
    lam = np.real(np.fft.fftn(cov))
    lam = np.maximum(lam, 0.0)
    noise = rng.standard_normal((pz, py, px))
    field = np.real(np.fft.ifftn(np.sqrt(lam) * np.fft.fftn(noise)))
    field = field[:nz, :ny, :nx]
    field = (field - field.mean()) / max(field.std(), 1e-12)
    return sigma * field


def correlation_pvt_rows(p_lo=10.0, p_hi=450.0, dp=20.0) -> np.ndarray:
    """Live-oil saturated table from polynomial correlations (bubble point 200 bar).

    Rs(p) = 0.4 p + 5e-4 p^2 (Rs(200) = 100 sm3/sm3), Bo = 1.05 + 1.5e-3 Rs,
    mu_o = 2.0 - 3e-3 p + 2e-6 p^2 cP, dry gas Bg = 1.2/p rm3/sm3,
    mu_g = 0.012 + 3e-5 p cP. One row per pressure, in the reference's PVTO
    row format (p, Rs, Bo, mu_o, Bg, mu_g), rs/pvt.rs:40-48.
    """
    p = np.arange(p_lo, p_hi + 1e-9, dp)
    rs = 0.4 * p + 5e-4 * p * p
    bo = 1.05 + 1.5e-3 * rs
    mu_o = 2.0 - 3e-3 * p + 2e-6 * p * p
    bg = 1.2 / p
    mu_g = 0.012 + 3e-5 * p
    return np.stack([p, rs, bo, mu_o, bg, mu_g], axis=1)


# --------------------------------------------------------------------------- decks
def waterflood_1d(n_steps: int = 150) -> Deck:
    """Two-phase 1D waterflood: 200x1x1, dx=1 m, dy=dz=10 m, phi=0.2, k=100 mD,
    mu_w=0.5, mu_o=2.0 cP, Corey 2/2 with Swc=Sor=0.1, Sw0=Swc, p0=300 bar,
    0.5 m3/d rate injector in cell 0, 250 bar BHP producer in cell 199,
    fixed 1-day report steps."""
    nx = 200
    n = nx
    calls = [
        ("set_cell_dimensions", (1.0, 10.0, 10.0)),
        ("set_permeability_field", (np.full(n, 100.0), np.full(n, 100.0), np.full(n, 100.0))),
        ("set_fluid_properties", (2.0, 0.5)),
        ("set_rel_perm_props", (0.1, 0.1, 2.0, 2.0, 1.0, 1.0)),
        ("set_capillary_params", (0.0, 2.0)),
        ("set_gravity_enabled", (False,)),
        ("set_initial_pressure", (300.0,)),
        ("set_initial_saturation", (0.1,)),
        ("set_well_control_modes", ("rate", "pressure")),
        ("set_target_well_rates", (0.5, 0.0)),
        ("add_well", (0, 0, 0, 400.0, 0.1, 0.0, True)),
        ("add_well", (nx - 1, 0, 0, 250.0, 0.1, 0.0, False)),
    ]
    return Deck("waterflood_1d_200", nx, 1, 1, 0.2, calls, [1.0] * n_steps)


def _three_phase_common(calls, rows):
    calls += [
        ("set_three_phase_mode_enabled", (True,)),
        ("set_fluid_properties", (1.5, 0.5)),
        ("set_fluid_compressibilities", (1e-5, 3e-6)),
        ("set_fluid_densities", (850.0, 1000.0)),
        ("set_gas_fluid_properties", (0.02, 1e-4, 0.9)),
        ("set_rel_perm_props", (0.15, 0.15, 2.0, 2.0, 1.0, 1.0)),
        ("set_three_phase_rel_perm_props", (0.15, 0.15, 0.05, 0.05, 0.1, 2.0, 2.0, 3.0, 0.6, 1.0, 0.8)),
        ("set_pvt_table", (rows,)),
    ]


def black_oil_3d(nx=64, ny=64, nz=16, dx=30.0, dy=30.0, dz=5.0, sigma=1.0, corr=8.0, seed=0x5EED0002,
                 n_steps=60, dt=5.0, gravity=False, leverett=False, five_wells=False, p_init=250.0,
                 name=None) -> Deck:
    """Three-phase black-oil FIM deck.

    ln k ~ GRF(mean ln 100 mD, sigma, exponential covariance, correlation
    length `corr` cells), kz = 0.1 kh; porosity U[0.1, 0.3] (seed + 0x1000);
    live oil from `correlation_pvt_rows`; Stone II from Corey 2/2/3.
    Initial state undersaturated at Rs = Rs_sat(200 bar), Sw = 0.2.
    Quarter five-spot (gas injector at (0,0), producer at (nx-1,ny-1), all
    layers) or, with `five_wells`, a water injector, a gas injector and three
    BHP producers, all completed in every layer.
    """
    n = nx * ny * nz
    lnk = np.log(100.0) + gaussian_random_field((nz, ny, nx), corr, sigma, seed).ravel()
    kh = np.exp(lnk)
    kz = 0.1 * kh
    rng = np.random.default_rng(seed + 0x1000)
    phi = rng.uniform(0.1, 0.3, size=n)
    rows = correlation_pvt_rows()
    rs_b = 0.4 * 200.0 + 5e-4 * 200.0 ** 2
    calls: List[Tuple[str, tuple]] = [
        ("set_cell_dimensions", (dx, dy, dz)),
        ("set_porosity_field", (phi,)),
        ("set_permeability_field", (kh, kh.copy(), kz)),
        ("set_rock_properties", (0.0, 2000.0, 1.0, 1.0)),
    ]
    _three_phase_common(calls, rows)
    calls += [("set_capillary_params", (2.0 if leverett else 0.0, 2.0)),
              ("set_gas_oil_capillary_params", (0.0, 2.0)),
              ("set_gravity_enabled", (bool(gravity),))]
    if leverett:
        # Leverett-J: Pc scales with sqrt(phi/k) relative to (0.2, 100 mD).
        calls.append(("set_capillary_scale_field", (np.sqrt((phi / kh) / (0.2 / 100.0)),)))
    # initial state: hydrostatic oil gradient with gravity, else uniform
    k_idx = np.repeat(np.arange(nz), nx * ny)
    depth = 2000.0 + (k_idx + 0.5) * dz
    p0 = p_init + (850.0 * 9.80665 * (depth - depth[0]) * 1e-5 if gravity else 0.0)
    p0 = np.asarray(p0, dtype=np.float64) * np.ones(n)
    calls.append(("set_initial_state_arrays", (p0, np.full(n, 0.2), np.zeros(n), np.full(n, rs_b))))
    calls.append(("set_injected_fluid", ("gas",)))
    calls.append(("set_well_bhp_limits", (50.0, 600.0)))
    # injection targets per completion: ~2% of a completed cell's pore volume per day at
    # reservoir conditions (gas Bg ~ 5e-3 rm3/sm3 at 250 bar)
    cell_pv = dx * dy * dz * 0.2
    gas_rate = 0.02 * cell_pv / 5e-3 * nz
    water_rate = 0.02 * cell_pv * nz
    if not five_wells:
        for k in range(nz):
            calls.append(("add_well_with_id", (0, 0, k, 400.0, 0.1, 0.0, True, "GINJ")))
        for k in range(nz):
            calls.append(("add_well_with_id", (nx - 1, ny - 1, k, 150.0, 0.1, 0.0, False, "PROD")))
        calls.append(("set_well_schedule", ("GINJ", "rate", np.nan, gas_rate, 450.0, True)))
        calls.append(("set_well_schedule", ("PROD", "pressure", np.nan, np.nan, 150.0, True)))
    else:
        layout = [("WINJ", 0, 0, True), ("GINJ", nx - 1, ny - 1, True), ("P1", nx - 1, 0, False),
                  ("P2", 0, ny - 1, False), ("P3", nx // 2, ny // 2, False)]
        for wid, i, j, inj in layout:
            for k in range(nz):
                calls.append(("add_well_with_id", (i, j, k, 400.0 if inj else 170.0, 0.1, 0.0, inj, wid)))
        calls.append(("set_well_schedule", ("WINJ", "rate", np.nan, water_rate, 500.0, True)))
        calls.append(("set_well_schedule", ("GINJ", "rate", np.nan, 1.5 * gas_rate, 500.0, True)))
        calls.append(("set_well_injected_fluid", ("WINJ", "water")))
        calls.append(("set_well_injected_fluid", ("GINJ", "gas")))
        for wid in ("P1", "P2", "P3"):
            calls.append(("set_well_schedule", (wid, "pressure", np.nan, np.nan, 170.0, True)))
    d = Deck(name or f"black_oil_{nx}x{ny}x{nz}", nx, ny, nz, 0.2, calls, [dt] * n_steps,
             info={"lnk_sha256": sha256_f64(lnk), "phi_sha256": sha256_f64(phi)})
    return d


def black_oil_65k(n_steps=60) -> Deck:
    """64x64x16 three-phase quarter five-spot, 30x30x5 m cells, 60 x 5 d."""
    return black_oil_3d(64, 64, 16, 30.0, 30.0, 5.0, sigma=1.0, seed=0x5EED0002, n_steps=n_steps, dt=5.0,
                        name="black_oil_64x64x16")


def black_oil_1m(n_steps=30) -> Deck:
    """128x128x64 three-phase, sigma 1.5, gravity, Leverett-J Pc, five
    multi-layer wells, 30 x 10 d."""
    return black_oil_3d(128, 128, 64, 30.0, 30.0, 5.0, sigma=1.5, seed=0x5EED0003, n_steps=n_steps, dt=10.0,
                        gravity=True, leverett=True, five_wells=True, p_init=250.0, name="black_oil_128x128x64")


def black_oil_8m(n_steps=4) -> Deck:
    """256x256x128 (8.4M cells): the 128x128x64 recipe at twice the resolution in every
    direction (15x15x2.5 m cells, the same wells by position), 10 d report steps. A
    scaling point for one GPU's HBM, not a parity case."""
    return black_oil_3d(256, 256, 128, 15.0, 15.0, 2.5, sigma=1.5, corr=16.0, seed=0x5EED0006, n_steps=n_steps,
                        dt=10.0, gravity=True, leverett=True, five_wells=True, p_init=250.0,
                        name="black_oil_256x256x128")


def black_oil_adjoint(n_steps=50) -> Deck:
    """48x48x8 three-phase with the 64x64x16 deck's PVT and relperm, 50 x 5 d."""
    return black_oil_3d(48, 48, 8, 30.0, 30.0, 5.0, sigma=1.0, seed=0x5EED0005, n_steps=n_steps, dt=5.0,
                        name="black_oil_48x48x8")


def without_redissolution(d: Deck) -> Deck:
    """The same deck with gas redissolution disabled (DRSDT 0, rs/frontend.rs:1208)."""
    return Deck(d.name + "_drsdt0", d.nx, d.ny, d.nz, d.porosity0, list(d.calls) + [("set_gas_redissolution_enabled", (False,))],
                list(d.schedule), info=dict(d.info))


def small_black_oil(nx=6, ny=5, nz=3, n_steps=3, dt=2.0, gravity=True, leverett=True, five_wells=False,
                    seed=7) -> Deck:
    """Small three-phase deck for parity tests (same recipe as the large decks)."""
    return black_oil_3d(nx, ny, nz, 30.0, 30.0, 5.0, sigma=1.0, corr=2.0, seed=seed, n_steps=n_steps, dt=dt,
                        gravity=gravity, leverett=leverett, five_wells=five_wells,
                        name=f"small_black_oil_{nx}x{ny}x{nz}")


def two_phase_member(m: int, n_steps=40, dt=2.0) -> Deck:
    """Ensemble member m: two-phase oil-water, nx,ny ~ U{32..96}, nz ~ U{4..8},
    sigma ~ U[0.5, 2], mu_o/mu_w ~ U[0.5, 50], Corey exponents ~ U[1.5, 4],
    injector and producer in opposite corners (all layers)."""
    seed = 0x5EED0004 + m
    rng = np.random.default_rng(seed)
    nx, ny = (int(v) for v in rng.integers(32, 97, size=2))
    nz = int(rng.integers(4, 9))
    sigma = float(rng.uniform(0.5, 2.0))
    ratio = float(rng.uniform(0.5, 50.0))
    nw, no = (float(v) for v in rng.uniform(1.5, 4.0, size=2))
    n = nx * ny * nz
    lnk = np.log(100.0) + gaussian_random_field((nz, ny, nx), 8.0, sigma, seed).ravel()
    kh = np.exp(lnk)
    phi = np.random.default_rng(seed + 0x1000).uniform(0.1, 0.3, size=n)
    mu_w = 0.5
    calls = [
        ("set_cell_dimensions", (20.0, 20.0, 4.0)),
        ("set_porosity_field", (phi,)),
        ("set_permeability_field", (kh, kh.copy(), 0.1 * kh)),
        ("set_fluid_properties", (mu_w * ratio, mu_w)),
        ("set_rel_perm_props", (0.1, 0.1, nw, no, 1.0, 1.0)),
        ("set_capillary_params", (0.0, 2.0)),
        ("set_initial_pressure", (300.0,)),
        ("set_initial_saturation", (0.1,)),
        ("set_well_bhp_limits", (100.0, 500.0)),
    ]
    for k in range(nz):
        calls.append(("add_well_with_id", (0, 0, k, 450.0, 0.1, 0.0, True, "INJ")))
        calls.append(("add_well_with_id", (nx - 1, ny - 1, k, 250.0, 0.1, 0.0, False, "PROD")))
    calls.append(("set_well_schedule", ("INJ", "rate", 40.0 * nz, np.nan, 500.0, True)))
    calls.append(("set_well_schedule", ("PROD", "pressure", np.nan, np.nan, 250.0, True)))
    return Deck(f"member_{m}", nx, ny, nz, 0.2, calls, [dt] * n_steps,
                info=dict(sigma=sigma, mobility_ratio=ratio, n_w=nw, n_o=no, lnk_sha256=sha256_f64(lnk),
                          phi_sha256=sha256_f64(phi)))


def small_waterflood(nx=8, ny=6, nz=3, n_steps=3, dt=1.0, rate=True, pc=True, seed=3) -> Deck:
    """Small two-phase 3D deck for parity tests."""
    n = nx * ny * nz
    rng = np.random.default_rng(seed)
    kh = np.exp(np.log(100.0) + 0.5 * rng.standard_normal(n))
    phi = rng.uniform(0.15, 0.25, size=n)
    calls = [
        ("set_cell_dimensions", (10.0, 10.0, 2.0)),
        ("set_porosity_field", (phi,)),
        ("set_permeability_field", (kh, kh.copy(), 0.1 * kh)),
        ("set_fluid_properties", (2.0, 0.5)),
        ("set_rel_perm_props", (0.1, 0.1, 2.0, 2.0, 1.0, 1.0)),
        ("set_capillary_params", (1.0 if pc else 0.0, 2.0)),
        ("set_gravity_enabled", (True,)),
        ("set_initial_pressure", (300.0,)),
        ("set_initial_saturation", (0.2,)),
    ]
    if rate:
        calls += [("set_well_control_modes", ("rate", "pressure")), ("set_target_well_rates", (20.0, 0.0))]
    for k in range(nz):
        calls.append(("add_well_with_id", (0, 0, k, 350.0, 0.1, 0.0, True, "I")))
        calls.append(("add_well_with_id", (nx - 1, ny - 1, k, 250.0, 0.1, 0.0, False, "P")))
    return Deck(f"small_waterflood_{nx}x{ny}x{nz}", nx, ny, nz, 0.2, calls, [dt] * n_steps)


# SPE1 Case 1 (Odeh, 1981) as the reference builds it (rs/lib.rs:286-609 make_spe1_like_grid_sim,
# rs/tests/spe1_acceptance.rs:124-141): 3048 m x 3048 m, three layers, live oil with gas
# injection, the published SWOF/SGOF tables. Table rows are the reference's values.
SPE1_PVT_ROWS = [
    (1.01, 0.18, 1.062, 1.040, 0.9361, 0.0080),
    (18.25, 16.12, 1.150, 0.975, 0.0679, 0.0096),
    (35.49, 32.06, 1.207, 0.910, 0.0352, 0.0112),
    (69.96, 66.08, 1.295, 0.830, 0.0179, 0.0140),
    (138.91, 113.29, 1.435, 0.695, 0.00906, 0.0189),
    (173.38, 138.03, 1.500, 0.641, 0.00727, 0.0208),
    (207.85, 165.64, 1.565, 0.594, 0.00607, 0.0228),
    (276.79, 226.20, 1.695, 0.510, 0.00455, 0.0268),
    (345.73, 288.17, 1.827, 0.449, 0.00364, 0.0309),
]
SPE1_SWOF = [  # sw, krw, krow, pcow
    (0.12, 0.0, 1.0, 0.0), (0.18, 4.64876033057851e-8, 1.0, 0.0), (0.24, 1.86e-7, 0.997, 0.0),
    (0.3, 4.18388429752066e-7, 0.98, 0.0), (0.36, 7.43801652892562e-7, 0.7, 0.0),
    (0.42, 1.16219008264463e-6, 0.35, 0.0), (0.48, 1.67355371900826e-6, 0.2, 0.0),
    (0.54, 2.27789256198347e-6, 0.09, 0.0), (0.6, 2.97520661157025e-6, 0.021, 0.0),
    (0.66, 3.7654958677686e-6, 0.01, 0.0), (0.72, 4.64876033057851e-6, 0.001, 0.0),
    (0.78, 5.625e-6, 0.0001, 0.0), (0.84, 6.69421487603306e-6, 0.0, 0.0), (0.91, 8.05914256198347e-6, 0.0, 0.0),
    (1.0, 1e-5, 0.0, 0.0),
]
SPE1_SGOF = [  # sg, krg, krog, pcog
    (0.0, 0.0, 1.0, 0.0), (0.001, 0.0, 1.0, 0.0), (0.02, 0.0, 0.997, 0.0), (0.05, 0.005, 0.98, 0.0),
    (0.12, 0.025, 0.7, 0.0), (0.2, 0.075, 0.35, 0.0), (0.25, 0.125, 0.2, 0.0), (0.3, 0.19, 0.09, 0.0),
    (0.4, 0.41, 0.021, 0.0), (0.45, 0.6, 0.01, 0.0), (0.5, 0.72, 0.001, 0.0), (0.6, 0.87, 0.0001, 0.0),
    (0.7, 0.94, 0.0, 0.0), (0.85, 0.98, 0.0, 0.0), (0.88, 0.984, 0.0, 0.0),
]


def spe1(nx=10, report_days=365.0, max_dt_days=30.0) -> Deck:
    """SPE1 Case 1 at nx x nx x 3 with the producer in the far corner, stepped to every
    yearly checkpoint up to `report_days` in report steps of at most 30 days
    (rs/tests/spe1_acceptance.rs:170-184 step_to). The reference's stability caps
    (set_stability_params) are IMPES-only and left out."""
    rows = [{"sw": a, "krw": b, "krow": c, "pcow": d} for a, b, c, d in SPE1_SWOF]
    gas = [{"sg": a, "krg": b, "krog": c, "pcog": d} for a, b, c, d in SPE1_SGOF]
    calls = [
        ("set_cell_dimensions_per_layer", (3048.0 / nx, 3048.0 / nx, [6.096, 9.144, 15.24])),
        ("set_fluid_properties", (0.51, 0.318)),
        ("set_fluid_compressibilities", (2.06e-4, 4.67e-5)),
        ("set_pvt_table", (np.array(SPE1_PVT_ROWS),)),
        ("set_initial_rs", (226.197,)),
        ("set_rock_properties", (4.35e-5, 2560.0, 1.695, 1.038)),
        ("set_fluid_densities", (860.0, 1033.0)),
        ("set_initial_pressure", (331.0,)),
        ("set_initial_saturation", (0.12,)),
        ("set_capillary_params", (0.0, 2.0)),
        ("set_gravity_enabled", (True,)),
        ("set_three_phase_rel_perm_props", (0.12, 0.12, 0.04, 0.04, 0.18, 2.0, 2.5, 1.5, 1e-5, 1.0, 0.984)),
        ("set_three_phase_scal_tables", ({"swof": rows, "sgof": gas},)),
        ("set_gas_fluid_properties", (0.027, 1e-4, 0.854)),
        ("set_three_phase_mode_enabled", (True,)),
        ("set_injected_fluid", ("gas",)),
        ("set_gas_redissolution_enabled", (False,)),
        ("set_well_control_modes", ("rate", "rate")),
        ("set_target_well_rates", (12_000.0, 5_400.0)),
        ("set_target_well_surface_rates", (2_831_680.0, 3_179.74)),
        ("set_well_bhp_limits", (69.0, 621.0)),
        ("set_permeability_per_layer", ([500.0, 50.0, 200.0], [500.0, 50.0, 200.0], [500.0, 50.0, 200.0])),
        ("add_well", (nx - 1, nx - 1, 2, 69.0, 0.0762, 0.0, False)),
        ("add_well", (0, 0, 0, 621.0, 0.0762, 0.0, True)),
    ]
    sched = []
    t = 0.0
    checkpoints = [365.0 * k for k in range(1, int(report_days // 365.0) + 1)]
    if not checkpoints or checkpoints[-1] < report_days - 1e-9:
        checkpoints.append(report_days)
    for target in checkpoints:
        while t < target - 1e-9:
            dt = min(max_dt_days, target - t)
            sched.append(dt)
            t += dt
    return Deck(f"spe1_{nx}x{nx}x3", nx, nx, 3, 0.3, calls, sched, info={"checkpoints": checkpoints})


# Buckley-Leverett displacement cases of rs/tests/buckley.rs:27-66 (a 24-cell 1D column between
# a 500-bar injector and a 100-bar producer, uniform 2000 mD, Corey relperms, Pc off)
BUCKLEY_CASES = {
    "A": dict(dt_days=0.5, s_wc=0.1, s_or=0.1, n_w=2.0, n_o=2.0, mu_w=0.5, mu_o=1.0, rel_tol=0.25),
    "B": dict(dt_days=0.25, s_wc=0.15, s_or=0.15, n_w=2.2, n_o=2.0, mu_w=0.6, mu_o=1.4, rel_tol=0.30),
}


def buckley_leverett(case="A", nx=24, max_steps=4000) -> Deck:
    """rs/tests/buckley.rs:141-162 build_buckley_simulator on the FIM path: new(nx, 1, 1, 0.2),
    the case's relperms and viscosities, Sw0 = Swc, k = 2000 mD (kz = 200 mD: the seeded
    random field with equal bounds), Pc off, BHP injector in cell 0 and producer in the last
    cell. The reference's stability caps are IMPES-only and left out."""
    c = BUCKLEY_CASES[case]
    n = nx
    calls = [
        ("set_rel_perm_props", (c["s_wc"], c["s_or"], c["n_w"], c["n_o"], 1.0, 1.0)),
        ("set_initial_saturation", (c["s_wc"],)),
        ("set_permeability_field", (np.full(n, 2000.0), np.full(n, 2000.0), np.full(n, 200.0))),
        ("set_capillary_params", (0.0, 2.0)),
        ("set_fluid_properties", (c["mu_o"], c["mu_w"])),
        ("add_well", (0, 0, 0, 500.0, 0.1, 0.0, True)),
        ("add_well", (nx - 1, 0, 0, 100.0, 0.1, 0.0, False)),
    ]
    return Deck(f"buckley_leverett_{case}_{nx}", nx, 1, 1, 0.2, calls, [c["dt_days"]] * max_steps, info=dict(c))
