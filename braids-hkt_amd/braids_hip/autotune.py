"""Launch autotuner for a resident river solver.

The launch-shape knobs of the library (braids_tune: the XCD-aware workgroup
order, and on FP32 solvers the fused face-and-update launch) do not change
any result, only the time per substep, and which is faster depends on the
mesh (its size against the Infinity Cache, its rotors, its cell order) and
on the device.  tune_river times every combination on the solver's own
mesh and state: it saves the resident state (braids_snapshot), runs a few
accepted substeps per candidate from that same state, restores it, and
keeps the fastest.  The solver's clock and step hint are restored too, so a
tuned run continues exactly as an untuned one would.
"""
import time

from . import _lib


def candidates(solver):
    out = [dict(remap=r, river_fused=0) for r in (1, 0)]
    if solver.precision == 32:
        out += [dict(remap=r, river_fused=1) for r in (1, 0)]
    return out


def tune_river(solver, substeps=12, repeats=2):
    """Pick the fastest launch shape for solver.river_run; returns
    dict(choice=..., ms_per_substep={shape: ms}).  The solver must be
    configured for the river path."""
    L, h = solver.L, solver.h
    t0, hint0 = solver.time, getattr(solver, 'river_hint', None)
    _lib.check(L.braids_snapshot(h), 'braids_snapshot')
    timings = {}
    for cand in candidates(solver):
        try:
            for k, v in cand.items():
                solver.tune(k, v)
        except RuntimeError:
            continue                      # e.g. the mesh exceeds the fused tables
        best = None
        for _ in range(repeats):
            _lib.check(L.braids_restore(h), 'braids_restore')
            solver.time, solver.river_hint = t0, hint0
            solver.river_run(2)           # first launches of this shape
            a = time.perf_counter()
            acc, rej = solver.river_run(substeps)
            el = (time.perf_counter() - a) / max(acc + rej, 1) * 1e3
            best = el if best is None else min(best, el)
        timings[tuple(sorted(cand.items()))] = best
    choice = dict(min(timings, key=timings.get))
    for k, v in choice.items():
        solver.tune(k, v)
    _lib.check(L.braids_restore(h), 'braids_restore')
    solver.time, solver.river_hint = t0, hint0
    return dict(choice=choice, ms_per_substep={', '.join(f'{k}={v}' for k, v in key): round(ms, 4)
                                               for key, ms in timings.items()})


__all__ = ['tune_river', 'candidates']
