"""Forward/backward interface of the discrete adjoint of the FIM step (an extension:
the reference has no adjoint).

    fn = FimAdjoint(deck, checkpoint_interval=10)
    J = fn.forward(lnk_shift, phi)        # runs the deck's report steps, recording
    g_lnk, g_phi = fn.backward()          # dJ/d(ln k shift), dJ/d(phi), per cell

The objective is the surface oil produced by the producers over the run. `lnk_shift`
is a per-cell log-multiplier applied to kx, ky and kz together (0 = the deck's
fields); `phi` replaces the deck's porosity. With `replay` set, forward() runs
exactly those substep lengths instead of the substep controller (used for finite
differences with a fixed schedule). The backward pass replays each checkpointed
segment on the device, recomputes the Jacobians and solves the transposed systems.
"""
from __future__ import annotations

import numpy as np

from .simulator import ReservoirSimulator


class FimAdjoint:
    def __init__(self, deck, checkpoint_interval=10, solver_options=None, replay=None):
        self.deck = deck
        self.interval = int(checkpoint_interval)
        self.solver_options = dict(solver_options or {})
        self.replay = None if replay is None else list(replay)
        fields = dict(deck.calls)
        self._k = [np.asarray(a, dtype=np.float64) for a in fields["set_permeability_field"]]
        self._phi = np.asarray(fields["set_porosity_field"][0], dtype=np.float64)
        self.sim = None
        self.report = None

    def close(self):
        if self.sim is not None:
            self.sim.close()
            self.sim = None

    def forward(self, lnk_shift=None, phi=None):
        """Run the schedule from the deck's initial state; returns (J, accepted substep lengths)."""
        self.close()
        s = self.deck.build(ReservoirSimulator)
        if lnk_shift is not None or phi is not None:
            m = np.exp(np.asarray(lnk_shift, dtype=np.float64)) if lnk_shift is not None else 1.0
            s.set_permeability_field(*(k * m for k in self._k))
            s.set_porosity_field(self._phi if phi is None else np.asarray(phi, dtype=np.float64))
        if self.solver_options:
            s.set_solver_options(**self.solver_options)
        s.adjoint_record(self.interval)
        if self.replay is None:
            for dt in self.deck.schedule:
                if s.step(dt) != 0:
                    raise RuntimeError("forward step failed: " + s.last_warning())
        else:
            for dt in self.replay:
                rc, rep = s.run_timestep(dt)
                if rc != 0:
                    raise RuntimeError(f"forward substep of {dt} d not converged after {rep.newton_iterations} "
                                       f"Newton iterations (scaled residual {rep.final_scaled_residual:.3e}) "
                                       + s.last_warning())
        self.sim = s
        J, _ = s.adjoint_objective()
        return J, s.accepted_dts()

    def backward(self):
        """dJ/d(ln k shift) and dJ/d(phi) of the last forward run."""
        if self.sim is None:
            raise RuntimeError("backward() needs a forward() first")
        g_lnk, g_phi, self.report = self.sim.adjoint_gradient()
        return g_lnk, g_phi
