"""Host side of the Exner bed update and the morphodynamic transaction.

Mirrors the bed part of ShorelinePhysics.finish (shoreline_solver.py:218-241):
after the hydraulic interval the bed is updated from the Meyer-Peter--Mueller
bedload of the new state; a rejected update (bed below the rock, solid
residual above 1e-10, increment above 5 % of the depth) raises and the owner
rolls the whole interval back (the reference's BendTrialRejected
transaction).  The bed stays an FP64 accumulator on the device."""
import ctypes

import numpy as np

from . import _lib
from .solver import HydraulicSolver, TrialRejected


class ExnerUpdate:
    def __init__(self, solver: HydraulicSolver, rock, d50=0.02, relative_density=1.65, porosity=0.4,
                 critical_shields=0.047, morfac=1.0, sediment_depth=0.001):
        self.solver = solver
        self.rock = np.ascontiguousarray(rock, dtype=np.float64).reshape(solver.nc)
        o = _lib.BraidsExnerOptions(float(d50), float(relative_density), float(porosity), float(critical_shields),
                                    float(morfac), float(sediment_depth), self.rock.ctypes.data)
        _lib.check(solver.L.braids_exner_configure(solver.h, ctypes.byref(o)), 'braids_exner_configure')
        self.last = None

    def finish(self, dt):
        """Bed update over dt from the resident state; raises TrialRejected
        (bed unchanged) on the reference's budget/positivity/increment test."""
        r = _lib.BraidsExnerResult()
        rc = _lib.check(self.solver.L.braids_exner_finish(self.solver.h, float(dt), ctypes.byref(r)),
                        'braids_exner_finish')
        self.last = dict(morphology=True, solid_volume_residual_m3=r.residual, solid_boundary_m3=r.boundary,
                         maximum_relative_bed_increment=r.change, donor_factor_min=r.donor_min)
        if rc == _lib.BRAIDS_REJECTED:
            raise TrialRejected('Shoreline sediment budget/positivity/increment rejection')
        return self.last

    def bed(self):
        z = np.empty(self.solver.nc)
        _lib.check(self.solver.L.braids_get_bed(self.solver.h, ctypes.c_void_p(z.ctypes.data)), 'braids_get_bed')
        return z


class Morphodynamics:
    """Outer transaction: hydraulic interval (RiverExplicit.advance) then the
    bed update; a rejection restores state, bed and ledgers and the interval
    is retried at half length."""

    def __init__(self, river, exner):
        self.river, self.exner = river, exner
        self.solver = river.solver

    def step(self, dt, time):
        L, h = self.solver.L, self.solver.h
        _lib.check(L.braids_snapshot(h), 'braids_snapshot')
        hint = self.river.step_hint
        try:
            report = self.river.advance(dt, time)
            report.update(self.exner.finish(dt))
            return report
        except TrialRejected:
            _lib.check(L.braids_restore(h), 'braids_restore')
            self.river.step_hint = hint
            self.solver.time = time
            raise

    def advance(self, dt, time, nsteps, min_dt=1e-6):
        """nsteps accepted outer steps of (at most) dt; returns (time, rejections)."""
        rejected = 0; t = time; k = 0; trial = dt
        while k < nsteps:
            try:
                self.step(trial, t)
                t += trial; k += 1; trial = min(dt, 2 * trial)
            except TrialRejected:
                rejected += 1; trial *= 0.5
                if trial < min_dt:
                    raise
        return t, rejected
