"""The oracle (plain-C FP64 restatement) against the reference solver built
from its own sources, and against what the mathematics fixes.

Fixtures (tests/golden/, written by tests/golden/make_golden.py from
oracle/_ref):
  wave_flux_ref.npz     1200 wave_flux evaluations of the reference Fortran
  trajectories_ref.npz  accepted states of the reference executable
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import swe
from braids_hip import mesh as M

GOLD = Path(__file__).resolve().parent / 'golden'


def test_wave_flux_matches_reference_fixture_bit_for_bit():
    d = np.load(GOLD / 'wave_flux_ref.npz')
    for x, y in zip(d['inputs'], d['outputs']):
        fl, fr, s, rm = swe.wave_flux(x[0:3], x[3:6], x[6], x[7], x[8:10])
        got = np.concatenate([fl, fr, [s, rm]])
        assert np.array_equal(got, y), (x, got - y)


@pytest.mark.skipif(not swe.reference_available(), reason='reference build (oracle/_ref) absent')
def test_wave_flux_matches_live_reference_random():
    ref = swe.Reference()
    rng = np.random.default_rng(123)
    for _ in range(3000):
        hl, hr = rng.uniform(0, 4, 2) * rng.integers(0, 2, 2)
        ul, ur, vl, vr = rng.normal(0, 4, 4)
        bl = rng.normal(); br = bl + rng.normal(0, 1)
        th = rng.uniform(0, 2 * np.pi); n = np.array([np.cos(th), np.sin(th)])
        L = np.array([hl, hl * ul, hl * vl]); R = np.array([hr, hr * ur, hr * vr])
        a = swe.wave_flux(L, R, bl, br, n); b = ref.wave_flux(L, R, bl, br, n)
        assert np.array_equal(np.concatenate([a[0], a[1], a[2:]]), np.concatenate([b[0], b[1], b[2:]]))


def _trajectory_mesh(name):
    if name.startswith('braided'):
        return M.braided_reach(nx=60, ny=12, length=300.0, width=60.0, amplitude=8.0, wavelength=150.0,
                               nbars=12, q=30.0, depth=1.0, bar_height=(0.3, 1.0), bar_sigma=(3.0, 10.0), bank=2.0)
    return M.plain_channel(nx=50, ny=6, length=200.0, width=24.0)


@pytest.mark.parametrize('name,scheme,tend', [('channel_s1', 1, 2.0), ('channel_s5', 5, 2.0), ('braided_s5', 5, 3.0)])
def test_trajectory_matches_reference_executable(name, scheme, tend):
    """Whole SSPRK(3,3) runs with dt control, retries, wet/dry fronts and both
    open-boundary modes reproduce the reference executable bit for bit."""
    d = np.load(GOLD / 'trajectories_ref.npz')
    m = _trajectory_mesh(name)
    o = swe.Oracle(m, scheme=scheme, cfl_target=0.9, dtmax=1.0)
    q, t, hint, n, retries, led = o.advance(m.q0, m.z, 10 ** 9, tend=tend)
    assert t == float(d[f'{name}_time'])
    assert n == int(d[f'{name}_steps'])
    assert np.array_equal(q, d[f'{name}_q'])
    assert np.array_equal(led, d[f'{name}_ledger'])


def test_lake_at_rest_stays_at_rest():
    """Reference gate: lake-at-rest max change < 1e-10 (hybrid_hydraulics.py:151-155),
    over a bumpy bed with emerged dry cells."""
    m = M.lake_at_rest()
    assert (m.q0[0] <= 1e-8).any() and (m.q0[0] > 0).any()
    o = swe.Oracle(m, scheme=5)
    q, t, hint, n, retries, led = o.advance(m.q0, m.z, 40)
    assert n == 40
    assert np.abs(q - m.q0).max() < 1e-10


@pytest.mark.parametrize('scheme', [1, 5])
def test_mass_balance_closes_against_boundary_ledger(scheme):
    m = M.plain_channel(nx=50, ny=6, length=200.0, width=24.0)
    o = swe.Oracle(m, scheme=scheme)
    m0 = o.mass(m.q0)
    q, t, hint, n, retries, led = o.advance(m.q0, m.z, 60)
    drift = abs(o.mass(q) - m0 - led.sum()) / m0
    assert drift < 1e-12


def _flux(l, r, bl, br, n):
    fl, fr, s, _ = swe.wave_flux(np.asarray(l, float), np.asarray(r, float), bl, br, np.asarray(n, float))
    return fl, fr, s


def test_wave_flux_mirror_symmetry():
    """Swapping sides and flipping the normal reverses every face flux."""
    rng = np.random.default_rng(5)
    for _ in range(500):
        hl, hr = rng.uniform(0.1, 3, 2)
        L = [hl, hl * rng.normal(), hl * rng.normal()]; R = [hr, hr * rng.normal(), hr * rng.normal()]
        bl, br = rng.normal(0, .3, 2)
        th = rng.uniform(0, 2 * np.pi); n = np.array([np.cos(th), np.sin(th)])
        fl, fr, s = _flux(L, R, bl, br, n)
        gl, gr, t = _flux(R, L, br, bl, -n)
        assert np.allclose(gl, -fr, rtol=1e-12, atol=1e-12)
        assert np.allclose(gr, -fl, rtol=1e-12, atol=1e-12)
        assert abs(s - t) <= 1e-12 * max(1, s)


def test_wave_flux_rotation_invariance():
    rng = np.random.default_rng(6)
    for _ in range(300):
        hl, hr = rng.uniform(0.1, 3, 2)
        L = np.array([hl, hl * rng.normal(), hl * rng.normal()]); R = np.array([hr, hr * rng.normal(), hr * rng.normal()])
        bl, br = rng.normal(0, .3, 2)
        n = np.array([1.0, 0.0])
        a = rng.uniform(0, 2 * np.pi); c, s_ = np.cos(a), np.sin(a)
        rot = np.array([[c, -s_], [s_, c]])
        fl, fr, s = _flux(L, R, bl, br, n)
        L2 = L.copy(); L2[1:] = rot @ L[1:]; R2 = R.copy(); R2[1:] = rot @ R[1:]
        gl, gr, t = _flux(L2, R2, bl, br, rot @ n)
        assert abs(gl[0] - fl[0]) <= 1e-11 * max(1, abs(fl[0]))
        assert np.allclose(gl[1:], rot @ fl[1:], rtol=1e-11, atol=1e-11)
        assert np.allclose(gr[1:], rot @ fr[1:], rtol=1e-11, atol=1e-11)


def test_still_water_over_a_step_is_hydrostatic():
    """Well-balanced known answer: equal free surface, zero velocity ->
    zero mass flux and each side's hydrostatic pressure 1/2 g h^2 n."""
    g = 9.81
    for hl, step in [(2.0, 0.5), (1.0, -0.3), (3.0, 1.2)]:
        hr = hl - step
        n = np.array([0.6, 0.8])
        fl, fr, s = _flux([hl, 0, 0], [hr, 0, 0], 0.0, step, n)
        assert abs(fl[0]) < 1e-13
        assert np.allclose(fl[1:], 0.5 * g * hl ** 2 * n, rtol=1e-13, atol=1e-13)
        assert np.allclose(fr[1:], 0.5 * g * hr ** 2 * n, rtol=1e-13, atol=1e-13)


def test_dry_step_reflects_like_a_wall():
    """A dry right cell whose bed tops the reflected surface acts as a wall:
    zero mass flux (wall mirroring of rpn2_geoclaw.f:170-187)."""
    fl, fr, s = _flux([1.0, 0.5, 0.0], [0.0, 0.0, 0.0], 0.0, 5.0, np.array([1.0, 0.0]))
    assert fl[0] == 0.0 and fr[0] == 0.0


def test_riemanntype_two_rarefaction_closed_form():
    """Two receding states: hm = (uL - uR + 2(cL + cR))^2 / (16 g)."""
    g = 9.81
    hl, hr, ul, ur = 2.0, 1.5, -1.0, 1.0
    hm, s1m, s2m, r1, r2 = swe.riemanntype(hl, hr, ul, ur)
    expect = (max(0.0, -(ur - ul) + 2 * (np.sqrt(g * hl) + np.sqrt(g * hr)))) ** 2 / (16 * g)
    assert r1 == 1 and r2 == 1
    assert abs(hm - expect) <= 1e-14 * expect


def test_manning_source_closed_form():
    """source (cutcell_solver.f90:14596-14598): hu / (1 + dt g n^2 |u| / h^(4/3))."""
    m = M.lake_at_rest(nx=2, ny=1)
    o = swe.Oracle(m, scheme=1)
    q = np.array([[1.5, 0.8, 1e-9, 2.0], [1.2, -0.4, 3e-9, 0.0], [0.3, 0.1, 0.0, 5.0]], dtype=float).reshape(3, 4)
    dt, g, n = 0.3, 9.81, m.manning[0]
    qq = q.copy()
    o.L.om_manning(o.ctx, qq.ctypes.data, dt)
    for c in range(4):
        h = q[0, c]
        if h <= 1e-8:
            assert qq[1, c] == 0 and qq[2, c] == 0
            continue
        sp = np.hypot(q[1, c], q[2, c]) / h
        d = 1 + dt * g * n ** 2 * sp / h ** (4 / 3)
        assert np.allclose(qq[1:, c], q[1:, c] / d, rtol=1e-15)


def test_dam_break_matches_ritter():
    """Dry-bed dam break against the Ritter solution (reference gate:
    analytic depth L1 < 0.02, hybrid_hydraulics.py:151-155)."""
    nx, L, h0 = 400, 200.0, 1.0
    m = M.lake_at_rest(nx=nx, ny=2, dx=L / nx, dy=1.0, stage=0.0, bumps=False)
    m.z[:] = 0.0
    m.manning[:] = 0.0          # frictionless, as the Ritter solution assumes
    x = m.xy[0]
    m.q0[0] = np.where(x < 100.0, h0, 0.0)
    o = swe.Oracle(m, scheme=5, cfl_target=0.9, dtmax=0.1)
    T = 8.0
    q, t, hint, n, retries, led = o.advance(m.q0, m.z, 10 ** 9, tend=T)
    g = 9.81; c0 = np.sqrt(g * h0); xi = (x - 100.0) / T
    exact = np.where(xi <= -c0, h0, np.where(xi >= 2 * c0, 0.0, (2 * c0 - xi) ** 2 / (9 * g)))
    l1 = np.sum(m.area * np.abs(q[0] - exact)) / np.sum(m.area * exact)
    assert l1 < 0.02, l1
    assert q[0].min() >= 0.0
