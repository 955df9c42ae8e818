"""Pinning of the river SSPRK(2,2) oracle (oracle/river_oracle.c).

* The four river operators match the reference's own compiled headers
  (river_explicit.h, river_midpoint.h) bit for bit: on the committed golden
  vectors always, and live when oracle/_ref is built.
* numpy.interp (the reference's table lookup, river_turbine_thetis.py:166-179)
  is matched bit for bit by the restatement, including knots and ends.
* The Strang drag reduces to the Manning closed form
  hu / (1 + dt g n^2 |u| / h^(4/3)) for the Manning law.
* The Thetis sink reproduces the actuator-disc thrust (1/2) A Ct U_inf^2 with
  the Kramer-Piggott induction alpha = (1 + sqrt(1 - A Ct / (D H)))/2 on a
  uniform stream, removes exactly the power it reports, and gates as the
  reference does (depth, bed, dry support, no flow, cut-in).
* The footprint quadrature integrates the bump to its closed form and assigns
  exactly the cells that overlap the rotor square.
* The river step conserves mass against its boundary ledger and keeps a lake
  at rest.
"""
from pathlib import Path

import numpy as np
import pytest

from braids_hip import mesh as M
from oracle import river as OR

GOLD = Path(__file__).resolve().parent / 'golden' / 'river_ops_ref.npz'
TABLE = dict(mode='thetis_speed_table', model_name='test rotor', performance=[
    dict(speed_m_s=s, thrust_coefficient=c, power_coefficient=p) for s, c, p in
    zip([0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 3.5], [0.8, 0.85, 0.85, 0.8, 0.7, 0.6, 0.5],
        [0.3, 0.4, 0.42, 0.42, 0.4, 0.35, 0.3])])


def rotor(X, Y, hub, control=TABLE, D=5.0):
    return dict(name=f'T{X:g}', X=X, Y=Y, D=D, hub_elevation_m=hub, minimum_bed_clearance_m=0.5,
                minimum_surface_clearance_m=0.5, control=control)


def flowing_channel(u=1.5, depth=8.0):
    m = M.plain_channel(nx=100, ny=20, length=200.0, width=40.0, depth=depth, q=u * depth * 40.0)
    m.q0[1] = m.q0[0] * u
    return m


# -- operators ---------------------------------------------------------------
@pytest.mark.parametrize('k', range(4))
def test_operators_match_reference_golden(k):
    from tests.golden.make_golden_river import case_mesh, run_ops
    g = np.load(GOLD)
    x = {n: g[f'c{k}_{n}'] for n in ('q', 'water', 'speed', 'trace', 'mask', 'base', 'rate0')}
    x['dt'] = float(g[f'c{k}_dt'])
    y = run_ops(OR.oracle_ops(), case_mesh(), x)
    for name in ('frequency', 'theta', 'water_out', 'correction', 'status', 'rate'):
        assert np.array_equal(y[name], g[f'c{k}_{name}']), name
    assert g[f'c{k}_theta'].min() < 1.0          # the cases exercise the limiter


def test_operators_match_live_reference():
    ops = OR.reference_ops()
    if ops is None:
        pytest.skip('oracle/_ref river operators not built here')
    from tests.golden.make_golden_river import case_inputs, run_ops
    m = M.braided_reach(nx=20, ny=6, length=100.0, width=30.0, nbars=4)
    for seed in range(20):
        x = case_inputs(m, 100 + seed)
        a, b = run_ops(ops, m, x), run_ops(OR.oracle_ops(), m, x)
        for name in a:
            assert np.array_equal(a[name], b[name]), (seed, name)


# -- table lookup ---------------------------------------------------------------
def test_interp_is_numpy_interp():
    rng = np.random.default_rng(4)
    xp = np.cumsum(rng.uniform(0.1, 1.0, 9)); fp = rng.normal(0, 1, 9)
    x = np.concatenate([rng.uniform(xp[0] - 1, xp[-1] + 1, 2000), xp, [xp[0] - 1e-300, xp[-1] + 0.0]])
    assert np.array_equal(OR.interp(x, xp, fp), np.interp(x, xp, fp))


def test_lookup_strict_upper_endpoint():
    xp = np.array([0.5, 1.0, 2.0]); fp = np.array([0.7, 0.8, 0.6])
    got = OR.lookup([0.49, 0.5, 0.75, 1.0, 1.999, 2.0, 3.0], xp, fp)
    assert got[0] == 0.0 and got[1] == 0.7 and got[3] == 0.8
    assert got[2] == np.interp(0.75, xp, fp) and got[4] == np.interp(1.999, xp, fp)
    assert got[5] == 0.0 and got[6] == 0.0     # x >= x_N is outside the table


# -- drag ------------------------------------------------------------------------
def test_drag_is_the_manning_closed_form():
    rng = np.random.default_rng(5)
    n = 500
    q = np.vstack([rng.uniform(1e-3, 5, n), rng.normal(0, 3, n), rng.normal(0, 3, n)])
    q[0, :20] = 1e-9                               # dry cells lose their momentum
    nm = rng.uniform(0.02, 0.05, n); dt = 0.37
    out, chezy = OR.cell_drag(q, dt, 2, nm)
    h = q[0]; sp = np.hypot(q[1], q[2]) / h
    expect = 1.0 / (1.0 + dt * 9.81 * nm ** 2 * sp / h ** (4.0 / 3.0))
    wet = h > 1e-8
    assert np.allclose(out[1, wet], q[1, wet] * expect[wet], rtol=2e-14, atol=0)
    assert np.all(out[1:, ~wet] == 0.0) and np.array_equal(out[0], q[0])
    assert np.allclose(chezy[wet], h[wet] ** (1 / 6) / nm[wet], rtol=1e-15)


def test_chezy_laws():
    assert OR.chezy(1, 2.0, 45.0) == 45.0
    assert np.isclose(OR.chezy(3, 2.0, 0.05), np.sqrt(9.81) / 0.41 * np.log(1 + 30 * 2.0 / (np.e * 0.05)), rtol=1e-14)
    assert np.isclose(OR.chezy(4, 2.0, 0.05), 18 * np.log10(12 * 2.0 / 0.05), rtol=1e-14)
    assert np.isclose(OR.chezy(4, 0.01, 0.05), 18 * np.log10(12 * 0.5), rtol=1e-14)


# -- footprint ---------------------------------------------------------------------
def _clip_area(tri, h):
    """Overlap area of a triangle with [-h, h]^2 by half-plane clipping (an
    independent Python implementation)."""
    poly = [tuple(p) for p in np.asarray(tri).reshape(3, 2)]
    for axis, sign in ((0, 1), (0, -1), (1, 1), (1, -1)):
        out = []
        for i in range(len(poly)):
            p, q = poly[i], poly[(i + 1) % len(poly)]
            fp, fq = h - sign * p[axis], h - sign * q[axis]
            if fp >= 0:
                out.append(p)
            if (fp >= 0) != (fq >= 0):
                t = fp / (fp - fq)
                out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
        poly = out
        if not poly:
            return 0.0
    a = 0.0
    for i in range(len(poly)):
        a += poly[i][0] * poly[(i + 1) % len(poly)][1] - poly[(i + 1) % len(poly)][0] * poly[i][1]
    return 0.5 * abs(a)


@pytest.mark.parametrize('X,Y,D', [(80.3, 20.1, 5.0), (100.0, 20.0, 5.0), (61.7, 9.9, 3.0)])
def test_footprint_integral_and_assignment(X, Y, D):
    m = flowing_channel()
    cand, val, area, order = OR.footprint_cells(m, X, Y, D)
    assert np.all(order >= 0)
    assert abs(val.sum() - OR.bump_total(D)) <= 1e-9 * OR.bump_total(D)
    assert abs(area.sum() - D * D) <= 1e-10 * D * D
    # exactly the cells overlapping the rotor square with positive area
    brute = np.array([_clip_area(m.verts[:, i] - np.array([X, Y] * 3), D / 2) for i in range(m.nc)])
    assert np.array_equal(np.sort(cand[val > 0]), np.flatnonzero(brute > 0))
    assert np.allclose(area, brute[cand], rtol=1e-12, atol=1e-14)


def test_density_normalisation_is_one():
    """sum(density) = bump integral / ((D/2)^2 1.45661) = 1 to the 6 digits of
    the reference constant (river_turbine_thetis.py:253)."""
    D = 5.0
    assert abs(OR.bump_total(D) / ((D / 2) ** 2 * 1.45661) - 1) < 5e-6


# -- Thetis sink -------------------------------------------------------------------
def test_uniform_stream_actuator_disc_thrust():
    """On a uniform stream u, depth H: F = sum K density u^2 = (1/2) A Ct
    (u/alpha)^2 sum(density), alpha = (1 + sqrt(1 - A Ct/(D H)))/2 (Kramer
    and Piggott 2016), and the sink removes exactly the reported power."""
    u, H, D = 1.5, 8.0, 5.0
    m = flowing_channel(u=u, depth=H)
    m.z[:] = 0.0
    m.q0[0] = H; m.q0[1] = H * u; m.q0[2] = 0.0
    o = OR.RiverOracle(m, rows=[rotor(80.3, 20.1, 0.5 + 2.5 + 0.5)])
    rate, diag, rc = o.turbines(m.q0, m.z)
    assert rc == 0 and diag[0, 0] == 1.0
    A = np.pi * D * D / 4
    ct = np.interp(u, [1.0, 1.5], [0.85, 0.85])
    alpha = 0.5 * (1 + np.sqrt(1 - A * ct / (D * H)))
    sdens = OR.bump_total(D) / ((D / 2) ** 2 * 1.45661)
    assert np.isclose(diag[0, 6], 0.5 * A * ct * (u / alpha) ** 2 * sdens, rtol=1e-12)
    assert np.isclose(diag[0, 9], u / alpha, rtol=1e-13)                 # incoming speed
    cp = np.interp(u / alpha, [1.5, 2.0], [0.42, 0.42])
    assert np.isclose(diag[0, 8], 0.5 * A * cp * (u / alpha) ** 3 * sdens, rtol=1e-12)
    # power removed from the flow = sum_c A_c (sink . u)
    removed = -np.sum(m.area * (rate[1] * u + rate[2] * 0.0))
    assert np.isclose(removed, diag[0, 7], rtol=1e-12)
    assert np.isclose(diag[0, 7], diag[0, 6] * u, rtol=1e-12)


@pytest.mark.parametrize('case,reason', [('shallow', 1), ('bed', 2), ('dry', 5), ('still', 7), ('slow', 8)])
def test_rotor_gating(case, reason):
    u, H = 1.5, 8.0
    m = flowing_channel(u=u, depth=H)
    m.z[:] = 0.0
    q = np.vstack([np.full(m.nc, H), np.full(m.nc, H * u), np.zeros(m.nc)])
    hub = 0.5 + 2.5 + 0.5
    if case == 'shallow':
        hub = H - 2.5 + 0.2                       # rotor top above the surface
    if case == 'bed':
        hub = 2.0                                 # rotor bottom below the bed
    if case == 'still':
        q[1] = 0.0
    if case == 'slow':
        q[1] = H * 0.3                            # below the first table speed
    o = OR.RiverOracle(m, rows=[rotor(80.3, 20.1, hub)])
    if case == 'dry':
        cells = o._tarr['foot_cell'][:1]
        q[:, cells] = 0.0
    rate, diag, rc = o.turbines(q, m.z)
    assert rc == 0 and diag[0, 0] == 0.0 and int(diag[0, 1]) == reason
    assert np.all(rate == 0.0)


def test_constant_ct_ideal_disk_power():
    ctl = dict(mode='constant_ct', thrust_coefficient=0.8, power_coefficient=None)
    u, H, D = 1.2, 8.0, 5.0
    m = flowing_channel(u=u, depth=H)
    m.z[:] = 0.0
    q = np.vstack([np.full(m.nc, H), np.full(m.nc, H * u), np.zeros(m.nc)])
    o = OR.RiverOracle(m, rows=[rotor(80.3, 20.1, 3.5, control=ctl)])
    _, diag, rc = o.turbines(q, m.z)
    A = np.pi * D * D / 4
    alpha = 0.5 * (1 + np.sqrt(1 - A * 0.8 / (D * H)))
    cp = 0.5 * 0.8 * (1 + np.sqrt(1 - 0.8))       # Thetis ideal disk
    sdens = OR.bump_total(D) / ((D / 2) ** 2 * 1.45661)
    assert rc == 0 and diag[0, 0] == 1.0
    assert np.isclose(diag[0, 8], 0.5 * A * cp * (u / alpha) ** 3 * sdens, rtol=1e-12)


def test_induction_domain_error():
    """A covered rotor whose footprint corners are shallow: A Ct/(D h) > 1 in
    those cells is the reference's ValueError (river_turbine_thetis.py:207-209)."""
    ctl = dict(mode='constant_ct', thrust_coefficient=0.95, power_coefficient=0.4)
    m = M.plain_channel(nx=120, ny=40, length=60.0, width=20.0, depth=8.0, q=100.0)
    X, Y = 30.1, 10.1
    m.z[:] = 0.0
    corner = (np.abs(m.xy[0] - X) > 1.6) & (np.abs(m.xy[0] - X) < 2.5) & (np.abs(m.xy[1] - Y) < 2.5)
    m.z[corner] = 5.0
    q = np.vstack([8.0 - m.z, 1.5 * (8.0 - m.z), np.zeros(m.nc)])
    o = OR.RiverOracle(m, rows=[dict(rotor(X, Y, 3.0, control=ctl), minimum_surface_clearance_m=0.0,
                                     minimum_bed_clearance_m=0.0)])
    _, diag, rc = o.turbines(q, m.z)
    assert rc == -1


# -- whole step -----------------------------------------------------------------------
def test_river_step_mass_balance_with_turbines():
    m = flowing_channel()
    o = OR.RiverOracle(m, rows=[rotor(80.3, 20.1, -0.08 + 3.3), rotor(120.7, 13.3, -0.12 + 3.3)])
    r = o.advance(m.q0, m.z, 1.0, 0.0)
    assert r['substeps'] > 0 and r['cfl_max'] <= 0.5 * (1 + 1e-12)
    drift = abs(o.hyd.mass(r['q']) - o.hyd.mass(m.q0) - r['ledger'].sum()) / o.hyd.mass(m.q0)
    assert drift <= 1e-12
    assert r['q'][0].min() >= 0


def test_river_lake_at_rest():
    m = M.lake_at_rest()
    o = OR.RiverOracle(m)
    r = o.advance(m.q0, m.z, 2.0, 0.0)
    assert r['substeps'] > 0
    assert np.abs(r['q'] - m.q0).max() < 1e-12


def test_river_rejection_halves_or_caps_the_hint():
    """A first trial far above the CFL limit is rejected and retried with
    min(dt/2, 0.98 cfl / peak) (river_explicit.py:136-139)."""
    m = flowing_channel()
    o = OR.RiverOracle(m)
    rc, _, _, peak = o.single(m.q0, m.z, 0.1, 0.0)
    assert rc == OR_CFL and 0.1 * peak > 0.5
    # a trial whose outflow exceeds a boundary cell's stock is rejected before
    # the CFL frequency (the limiter status comes first in RiverMidpoint.rate)
    rc, _, _, peak = o.single(m.q0, m.z, 1.0, 0.0)
    assert rc == OR_BOUNDARY and peak == 0.0
    o.hint = 0.1
    r = o.advance(m.q0, m.z, 0.2, 0.0)
    assert r['rejections'] >= 1 and r['cfl_max'] <= 0.5 * (1 + 1e-12)


OR_CFL = 2
OR_BOUNDARY = 3
