"""The boundary's error behaviour, from the reference's own API-contract tests
(src/lib/ressim/src/tests/runtime_api.rs:192-286): an invalid setter argument returns
RSF_ERR_INVALID with the reference's message in rsf_last_warning (a ValueError in the
Python mirror) and leaves the state unchanged. No GPU is needed: setters only touch the
host model."""
import math

import pytest

from ressim_amd import ReservoirSimulator


def completions(sim):
    """rsf_num_perforations: the reference's wells.len() (one entry per completion)."""
    return sim._l.rsf_num_perforations(sim._h)


def err_contains(call, text):
    with pytest.raises(ValueError) as e:
        call()
    assert text in str(e.value), str(e.value)


def test_multiple_wells_in_same_block_are_rejected_without_state_change():
    """runtime_api.rs:192-202."""
    sim = ReservoirSimulator(4, 1, 1, 0.2)
    sim.add_well(0, 0, 0, 600.0, 0.1, 0.0, True)
    err_contains(lambda: sim.add_well(0, 0, 0, 550.0, 0.1, 0.0, True), "same cell")
    assert completions(sim) == 1


def test_out_of_bounds_well_is_rejected_without_state_change():
    """runtime_api.rs:204-212."""
    sim = ReservoirSimulator(2, 2, 1, 0.2)
    err_contains(lambda: sim.add_well(2, 0, 0, 250.0, 0.1, 0.0, False), "out of bounds")
    assert completions(sim) == 0


def test_rejects_invalid_relperm_parameters():
    """runtime_api.rs:249-264, and the zero water endpoint of :266-271 is accepted."""
    sim = ReservoirSimulator(2, 2, 1, 0.2)
    err_contains(lambda: sim.set_rel_perm_props(0.6, 0.5, 2.0, 2.0, 1.0, 1.0), "must be < 1.0")
    err_contains(lambda: sim.set_rel_perm_props(0.1, 0.1, 0.0, 2.0, 1.0, 1.0), "must be positive")
    err_contains(lambda: sim.set_rel_perm_props(math.nan, 0.1, 2.0, 2.0, 1.0, 1.0), "finite numbers")
    sim.set_rel_perm_props(0.1, 0.1, 2.0, 2.0, 0.0, 1.0)


def test_rejects_invalid_density_and_capillary_inputs():
    """runtime_api.rs:273-286."""
    sim = ReservoirSimulator(2, 2, 1, 0.2)
    err_contains(lambda: sim.set_fluid_densities(-800.0, 1000.0), "must be positive")
    err_contains(lambda: sim.set_fluid_densities(800.0, math.nan), "finite numbers")
    err_contains(lambda: sim.set_capillary_params(-1.0, 2.0), "non-negative")
    err_contains(lambda: sim.set_capillary_params(5.0, 0.0), "positive")
    err_contains(lambda: sim.set_capillary_params(math.nan, 2.0), "finite numbers")


def test_rejects_invalid_fluid_rock_and_grid_inputs():
    """The same contract for the other scalar setters (rs/frontend.rs:621-847)."""
    sim = ReservoirSimulator(2, 2, 1, 0.2)
    err_contains(lambda: sim.set_cell_dimensions(10.0, 0.0, 1.0), "must be positive")
    err_contains(lambda: sim.set_cell_dimensions(10.0, math.inf, 1.0), "finite numbers")
    err_contains(lambda: sim.set_fluid_properties(0.0, 0.5), "must be positive")
    err_contains(lambda: sim.set_fluid_compressibilities(-1e-5, 1e-5), "non-negative")
    err_contains(lambda: sim.set_rock_properties(-1.0, 2000.0, 1.0, 1.0), "non-negative")
    err_contains(lambda: sim.set_rock_properties(0.0, 2000.0, 0.0, 1.0), "must be positive")
    err_contains(lambda: sim.add_well(0, 0, 0, math.nan, 0.1, 0.0, True), "BHP must be finite")
    err_contains(lambda: sim.add_well(0, 0, 0, 300.0, 0.0, 0.0, True), "Well radius")


def test_invalid_argument_without_a_reference_message_names_the_entry_point():
    sim = ReservoirSimulator(2, 2, 1, 0.2)
    err_contains(lambda: sim.set_porosity_field([0.2, 0.2]), "rsf_set_porosity_field: invalid argument")
    sim.set_fluid_properties(1.0, 0.5)  # a valid call leaves the last message alone
    assert "rsf_set_porosity_field" in sim.last_warning()


def test_ilu_sweep_setter_rejects_invalid_arguments():
    """rsf_set_ilu_sweep (an extension): mode 0 or 1, 1 to 256 grid lines per tile."""
    sim = ReservoirSimulator(4, 3, 2, 0.2)
    sim.set_ilu_sweep(0, 16, 16)
    sim.set_ilu_sweep(1, 256, 1)
    for args in ((2, 64, 64), (-1, 64, 64), (1, 0, 64), (1, 64, 257), (1, 257, 64)):
        with pytest.raises(ValueError):
            sim.set_ilu_sweep(*args)
