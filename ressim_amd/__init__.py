"""ressim_amd: MI355X-native fully implicit (FIM) Newton timestep of the
three-phase black-oil reservoir simulator, behind the C ABI in
include/ressim_fim.h.

    from ressim_amd import ReservoirSimulator, decks
    sim = decks.waterflood_1d().build(ReservoirSimulator)
    sim.step(1.0)
"""
from . import decks
from .simulator import ReservoirSimulator, default_solver_options, lib, library_path

__all__ = ["ReservoirSimulator", "decks", "default_solver_options", "lib", "library_path"]
