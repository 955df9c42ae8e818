"""Batched serving: a scenario's result is the same alone, in a mix and when
other scenarios are admitted while it runs, and equals the single-scenario
path (bit for bit in FP64; FP32 within 1e-12 relative, the two paths being
the same device bodies compiled in different kernels)."""
import numpy as np
import pytest

from braids_hip import mesh as M

pytestmark = pytest.mark.gpu

TABLE = M.THETIS_TABLE


def scenario(kind):
    if kind == 'channel':
        m = M.plain_channel(nx=100, ny=20, length=200.0, width=40.0, depth=8.0, q=480.0)
        m.q0[1] = m.q0[0] * 1.5
        rows = [dict(name='T1', X=80.3, Y=20.1, D=5.0, hub_elevation_m=-0.08 + 3.3, minimum_bed_clearance_m=0.5,
                     minimum_surface_clearance_m=0.5, control=TABLE)]
        return m, rows
    if kind == 'braided':
        return M.braided_reach(nx=60, ny=12, length=300.0, width=60.0, amplitude=8.0, wavelength=150.0, nbars=12,
                               q=30.0, depth=1.0, bar_height=(0.3, 1.0), bar_sigma=(3.0, 10.0), bank=2.0), []
    m = M.plain_channel(nx=40, ny=8, length=160.0, width=32.0, depth=3.0, q=150.0)
    m.q0[1] = m.q0[0] * 1.0
    return m, []


def solver(kind, precision):
    from braids_hip.solver import HydraulicSolver
    from braids_hip.river import RiverExplicit
    from braids_hip.turbines import ThetisArray
    m, rows = scenario(kind)
    s = HydraulicSolver(m, precision=precision, scheme=5)
    s.set_state(m.q0, prepare=False)
    RiverExplicit(s, cfl=.5)
    if rows:
        ThetisArray(m, rows).install(s)
    return s


def run_batch(kinds, precision, steps, late=None):
    """Run scenarios `kinds` in one batch; `late` = (kind, after) admits one
    more scenario after `after` substeps.  Returns {kind: (state, status)}."""
    from braids_hip.serve import Batch
    b = Batch(precision)
    sol = {k: solver(k, precision) for k in kinds}
    ids = {b.add(s, steps, 0.0, 0.01): k for k, s in sol.items()}
    out = {}
    if late:
        b.run(late[1])
        s = solver(late[0], precision)
        sol[late[0] + '*'] = s
        ids[b.add(s, steps, 0.0, 0.01)] = late[0] + '*'
    while ids:
        b.run(16)
        for st in b.poll():
            if st.done:
                k = ids.pop(st.id)
                fin = b.remove(st.id)
                out[k] = (sol[k].state(), fin.accepted, fin.rejections, sol[k].ledger())
    b.close()
    return out


@pytest.mark.parametrize('precision', [64, 32])
def test_batch_alone_matches_single_path(precision):
    steps = 60
    res = run_batch(['channel'], precision, steps)
    s = solver('channel', precision)
    acc, rej = s.river_run(steps, hint=0.01)
    qb, accb, rejb, ledb = res['channel']
    q = s.state()
    print(f'[serve] alone vs single precision={precision}: accepted {accb}/{acc} rejections {rejb}/{rej} '
          f'max|dq| {np.abs(qb - q).max():.3e}', flush=True)
    assert accb == acc == steps and rejb == rej
    if precision == 64:
        assert np.array_equal(qb, q)
        assert np.array_equal(ledb, s.ledger())
    else:
        assert np.abs(qb - q).max() <= 1e-12 * np.abs(q).max()


@pytest.mark.parametrize('precision', [64, 32])
def test_result_same_alone_and_in_the_mix(precision):
    steps = 50
    alone = {k: run_batch([k], precision, steps)[k] for k in ('channel', 'braided', 'small')}
    mix = run_batch(['channel', 'braided', 'small'], precision, steps)
    late = run_batch(['braided', 'small'], precision, steps, late=('channel', 17))
    for k in alone:
        qa, acc, rej, led = alone[k]
        for name, res in (('mix', mix[k]), ('late', late.get(k + '*', late.get(k)))):
            qm, accm, rejm, ledm = res
            print(f'[serve] {k} {name}: accepted {accm}/{acc} rejections {rejm}/{rej}', flush=True)
            assert accm == acc == steps and rejm == rej
            assert np.array_equal(qm, qa), (k, name, np.abs(qm - qa).max())
            assert np.array_equal(ledm, led)
