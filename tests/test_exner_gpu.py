"""GPU Exner bed update against the oracle (oracle/exner_oracle.c, pinned to
the reference MPM subroutine and bed operators in test_exner_oracle.py):
the bed increment to 1e-12 relative in FP64 and FP32 modes (the bed is FP64
in both), solid-volume conservation, the rock floor under donor limiting, the
rejection rule, and a morphodynamic sequence of river intervals plus bed
updates to 1e-10."""
import numpy as np
import pytest

from braids_hip import mesh as M

pytestmark = pytest.mark.gpu


def channel(u=2.0, depth=2.0, seed=3):
    m = M.plain_channel(nx=60, ny=6, length=120.0, width=12.0, depth=depth, q=u * depth * 12.0, n=0.03)
    rng = np.random.default_rng(seed)
    m.q0[1] = m.q0[0] * u * (1 + 0.3 * rng.standard_normal(m.nc))
    return m


def pair(m, rock, precision):
    from oracle import exner as OX
    from braids_hip.solver import HydraulicSolver
    from braids_hip.exner import ExnerUpdate
    s = HydraulicSolver(m, precision=precision, scheme=5)
    s.set_state(m.q0, prepare=False)
    return OX.ExnerOracle(m, rock), s, ExnerUpdate(s, rock)


@pytest.mark.parametrize('precision', [64, 32])
def test_bed_update_matches_oracle(precision):
    m = channel()
    rng = np.random.default_rng(5)
    rock = m.z - rng.uniform(0, 2e-4, m.nc)           # thin erodible layer: the donor limiter acts
    ox, s, ex = pair(m, rock, precision)
    r = ox.finish(m.q0, m.z, 0.5)
    g = ex.finish(0.5)
    zg = ex.bed()
    dz_o, dz_g = r['z'] - m.z, zg - m.z
    err = np.abs(dz_g - dz_o).max() / np.abs(dz_o).max()
    print(f'[exner] precision={precision} max rel increment err {err:.3e} donor {g["donor_factor_min"]:.4f}/{r["donor"]:.4f}',
          flush=True)
    assert r['ok'] and r['donor'] < 1.0
    assert err <= 1e-12, err
    assert np.all(zg >= rock - 1e-13)
    assert np.isclose(g['solid_boundary_m3'], r['boundary'], rtol=1e-12, atol=1e-18)
    assert abs(g['solid_volume_residual_m3']) <= 1e-10


def test_rejection_keeps_the_bed():
    from braids_hip.solver import TrialRejected
    m = M.plain_channel(nx=60, ny=6, length=120.0, width=12.0, depth=0.5, q=4.0 * 0.5 * 12.0, n=0.03)
    m.q0[1] = m.q0[0] * 4.0; m.q0[1, : m.nc // 2] *= 0.1
    ox, s, ex = pair(m, m.z - 5.0, 64)
    assert not ox.finish(m.q0, m.z, 500.0)['ok']
    with pytest.raises(TrialRejected):
        ex.finish(500.0)
    assert np.array_equal(ex.bed(), m.z)


def test_morphodynamic_sequence_matches_oracle():
    """Five outer steps of 0.5 s: river SSPRK(2,2) interval then the bed update."""
    from oracle import river as OR, exner as OX
    from braids_hip.river import RiverExplicit
    from braids_hip.exner import Morphodynamics
    m = channel()
    rock = m.z - 1.0
    ox, s, ex = pair(m, rock, 64)
    morph = Morphodynamics(RiverExplicit(s, cfl=.5), ex)
    orv = OR.RiverOracle(m, scheme=5, cfl=.5)
    q = m.q0.copy(); z = m.z.copy(); t = 0.0
    for k in range(5):
        ro = orv.advance(q, z, 0.5, t)
        fin = ox.finish(ro['q'], z, 0.5)
        assert fin['ok']
        morph.step(0.5, t)
        q, z, t = ro['q'], fin['z'], t + 0.5
    zg = ex.bed()
    err_z = np.abs((zg - m.z) - (z - m.z)).max() / np.abs(z - m.z).max()
    qe = [np.linalg.norm(s.state()[j] - q[j]) / np.linalg.norm(q[j]) for j in range(3)]
    print(f'[exner] morphodynamic 5 x 0.5 s: bed increment err {err_z:.3e}, state rel_l2 {qe}', flush=True)
    assert err_z <= 1e-10 and max(qe) <= 1e-10
