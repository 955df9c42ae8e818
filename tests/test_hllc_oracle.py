"""Pinning of the HLLC oracle (oracle/hllc_oracle.c, oracle/hllc.py).

* The eight operators (hllc.h, hllc_turbines.h, hllc_advanced.h) match the
  reference's own g++ build bit for bit: on the committed golden vectors
  always, and live when oracle/_ref is built.
* The HLLC flux is consistent (equal states give the physical flux) and
  mirror symmetric.
* The reconstruction recovers a linear field's gradient exactly where the
  limiter is inactive (this pins the inverse metric and the LS weights).
* The Hancock owner keeps a lake at rest (with dry islands), conserves mass
  exactly in a closed basin and against its boundary ledger in an open
  channel, and converges to the exact Stoker (wet) and Ritter (dry bed) dam
  breaks.
"""
from pathlib import Path

import numpy as np
import pytest
from scipy.optimize import brentq

from braids_hip import mesh as M
from oracle import hllc as OH
from tests.golden.make_golden_hllc import CASES, GEOM, case_mesh, chain_inputs, run_chain
from tests.test_river_oracle import flowing_channel, rotor

GOLD = Path(__file__).resolve().parent / 'golden' / 'hllc_ref.npz'


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize('k', range(len(CASES)))
def test_operators_match_reference_golden(k):
    d = np.load(GOLD)
    g = {n: d['g_' + n] for n in GEOM}
    pre = f'c{k}_in_'
    x = {n[len(pre):]: d[n] for n in d.files if n.startswith(pre)}
    y = run_chain(OH.oracle_ops(), g, x)
    for name, val in y.items():
        assert _same(val, d[f'c{k}_out_{name}']), name


def test_golden_cases_cover_the_branches():
    d = np.load(GOLD)
    th = np.concatenate([d[f'c{k}_out_theta'] for k in range(len(CASES))])
    assert ((th > 0) & (th < 1)).sum() > 10 and (th < 0).sum() > 0      # limiter and unfunded withdrawal
    assert sum(d[f'c{k}_out_pstatus'].sum() for k in range(len(CASES))) > 0
    assert any((d[f'c{k}_out_fl1'] != d[f'c{k}_out_fl0']).any() for k in range(len(CASES)))


def test_geometry_matches_golden():
    g = OH.geometry(case_mesh())
    d = np.load(GOLD)
    for n in GEOM:
        assert np.allclose(g[n], d['g_' + n], rtol=1e-13, atol=1e-15), n


def test_operators_match_live_reference():
    ref = OH.reference_ops()
    if ref is None:
        pytest.skip('oracle/_ref not built here')
    m = case_mesh(nx=14, ny=5)
    g = OH.geometry(m)
    for seed in range(30, 40):
        x = chain_inputs(m, seed, dt=[None, 0.8][seed % 2])
        a, b = run_chain(ref, g, x), run_chain(OH.oracle_ops(), g, x)
        for n in a:
            assert _same(a[n], b[n]), (seed, n)


# -- flux ---------------------------------------------------------------------
def test_flux_consistency():
    g = 9.81
    for h, u, v in ((1.0, 0.3, -0.2), (2.5, -1.1, 0.4), (0.3, 2.5, 0.0), (4.0, 0.0, 0.0)):
        out, s = OH.flux(h, u, v, h, u, v, g)
        want = np.array([h * u, h * u * u + .5 * g * h * h, h * u * v])
        assert np.allclose(out, want, rtol=1e-14, atol=1e-14)
        assert np.isclose(s, abs(u) + np.sqrt(g * h))


def test_flux_mirror_symmetry():
    rng = np.random.default_rng(4)
    for _ in range(200):
        hl, hr = rng.uniform(0, 3, 2)
        if rng.random() < 0.2:
            hl = 0.0
        ul, ur, vl, vr = rng.normal(0, 2, 4)
        a, sa = OH.flux(hl, ul, vl, hr, ur, vr)
        b, sb = OH.flux(hr, -ur, vr, hl, -ul, vl)
        assert np.allclose([a[0], a[1], a[2]], [-b[0], b[1], -b[2]], rtol=1e-12, atol=1e-12)
        assert np.isclose(sa, sb)


def test_dry_dry_face_carries_nothing():
    out, s = OH.flux(0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    assert not out.any() and s == 0.0


# -- reconstruction -------------------------------------------------------------
def test_linear_field_gradient_is_exact():
    m = M.lake_at_rest(nx=16, ny=8, bumps=False)
    x, y = m.xy
    a = np.array([0.01, -0.02, 0.05, 0.03, -0.04, 0.02, 0.004, -0.003])
    z = 0.1 + a[6] * x + a[7] * y
    eta = 3.0 + a[0] * x + a[1] * y
    u = 0.5 + a[2] * x + a[3] * y
    v = -0.2 + a[4] * x + a[5] * y
    h = eta - z
    q = np.ascontiguousarray(np.vstack([h, h * u, h * v]))
    g = OH.geometry(m)
    n, nf = m.nc, m.nf
    grad = np.zeros((8, n)); half = np.zeros((4, n)); qh = np.zeros((3, n)); st = np.zeros(n, np.int32)
    OH.oracle_ops()['hllc_reconstruct'](n, nf, 9.81, 1e-8, 0.0, q, np.ascontiguousarray(z), g['xy'], g['mid'],
                                        g['normal'], g['owner'], g['neighbour'], g['selected'], g['off'], g['faces'],
                                        g['inverse'], grad, half, qh, st)
    interior = np.ones(n, bool)
    for f in np.flatnonzero(m.neighbour < 0):
        interior[m.owner[f]] = False
    for f in np.flatnonzero(m.neighbour >= 0):     # and their neighbours (the limiter sees the walls' cells)
        if not interior[m.owner[f]] or not interior[m.neighbour[f]]:
            interior[m.owner[f]] = interior[m.neighbour[f]] = False
    assert interior.sum() > 100
    for k in range(8):
        assert np.allclose(grad[k, interior], a[k], rtol=1e-9, atol=1e-12), k
    # dt = 0: the predictor returns the cell values
    assert np.allclose(half[0], eta, rtol=1e-15) and np.allclose(half[1], u, rtol=1e-14, atol=1e-15)


def test_metric_inverse_is_a_true_inverse():
    m = case_mesh()
    g = OH.geometry(m)
    inv = g['inverse']
    # on every cell with >= 2 independent directions the metric is invertible
    det = inv[0] * inv[2] - inv[1] ** 2
    assert (det > 0).all()


# -- the owner ------------------------------------------------------------------
def closed_flume(nx=400, ny=2, length=100.0, width=1.0):
    m = M.structured_triangles(nx, ny, length / nx, width / ny, bed=lambda x, y: np.zeros_like(x),
                               manning=lambda x, y: 0.03, depth0=lambda x, y, z: np.ones_like(x), inflow_q=0.0,
                               outlet_stage=1.0, name='flume')
    m.bid[:] = -1; m.nb = 0; m.mode = np.zeros(0, np.int32)
    m.forcing_q = np.zeros(0); m.forcing_stage = np.zeros(0)
    m.build_incidence()
    return m


def dam_break(x, t, hl, hr, g=9.81):
    """Exact depth of the 1-D dam break (Stoker; Ritter when hr = 0)."""
    cl = np.sqrt(g * hl); xi = x / t
    if hr <= 0:
        return np.where(xi < -cl, hl, np.where(xi > 2 * cl, 0.0, (2 * cl - xi) ** 2 / (9 * g)))
    f = lambda hm: 2 * (cl - np.sqrt(g * hm)) - (hm - hr) * np.sqrt(g * (hm + hr) / (2 * hm * hr))
    hm = brentq(f, hr, hl)
    um = 2 * (cl - np.sqrt(g * hm)); cm = np.sqrt(g * hm); s = hm * um / (hm - hr)
    return np.where(xi < -cl, hl, np.where(xi < um - cm, (2 * cl - xi) ** 2 / (9 * g), np.where(xi < s, hm, hr)))


@pytest.mark.parametrize('hr,tol', [(0.5, 3e-3), (0.0, 4e-3)])
def test_dam_break_converges_to_exact(hr, tol):
    m = closed_flume()
    x = m.xy[0] - 50.0
    q = np.zeros((3, m.nc)); q[0] = np.where(x < 0, 2.0, hr)
    o = OH.HLLCOracle(m, cfl=.9, drag_mode=0)
    r = o.advance(q, m.z, 4.0)
    exact = dam_break(x, 4.0, 2.0, hr)
    err = np.sum(np.abs(r['q'][0] - exact) * m.area) / np.sum(exact * m.area)
    assert err < tol
    assert np.sum(r['q'][0] * m.area) == pytest.approx(np.sum(q[0] * m.area), rel=1e-14)
    assert r['q'][0].min() >= 0 and r['cfl_max'] <= .9 * (1 + 1e-12)


def test_lake_at_rest_with_dry_islands():
    """A basin with dry islands and banks stays at rest to roundoff over 80 s.
    (With a millimetre film on a cell beside a bank the scheme is not at rest
    to roundoff: the film's velocity noise grows to ~1e-3 over tens of
    seconds whichever rounding the LS metric has, so the islands here have
    no film shallower than 5 cm.)"""
    m = M.lake_at_rest(min_depth=0.05)
    assert (m.q0[0] <= 0).sum() > 20
    o = OH.HLLCOracle(m, cfl=.9)
    q, t = m.q0, 0.0
    for _ in range(4):
        r = o.advance(q, m.z, 20.0, t)
        q, t = r['q'], t + 20.0
    assert r['substeps'] > 10
    assert np.abs(q - m.q0).max() < 1e-13


def test_closed_basin_conserves_mass():
    m = M.lake_at_rest()
    x, y = m.xy
    q = m.q0.copy()
    q[0] = np.maximum(1.0 + 0.3 * np.exp(-((x - 40) ** 2 + (y - 12) ** 2) / 20) - m.z, 0.0)
    o = OH.HLLCOracle(m, cfl=.9, drag_mode=0)
    r = o.advance(q, m.z, 10.0)
    mass0 = np.sum(q[0] * m.area)
    assert abs(np.sum(r['q'][0] * m.area) - mass0) <= 1e-13 * mass0
    assert r['q'][0].min() >= 0


def test_open_channel_with_rotors_balances_the_ledger():
    m = flowing_channel()
    rows = [rotor(80.3, 20.1, -0.08 + 3.3), rotor(120.7, 13.3, -0.12 + 3.3)]
    o = OH.HLLCOracle(m, cfl=.9, rows=rows)
    o.hint = 0.05
    r = o.run(m.q0, m.z, 40)
    mass0 = np.sum(m.q0[0] * m.area)
    drift = abs(np.sum(r['q'][0] * m.area) - mass0 - r['ledger'].sum()) / mass0
    assert drift <= 1e-13
    assert r['q'][0].min() >= 0
    assert o.last_diag[:, 0].sum() == 2                # both rotors active
    # the rotors take momentum out of the stream
    o2 = OH.HLLCOracle(m, cfl=.9)
    o2.hint = 0.05
    r2 = o2.run(m.q0, m.z, 40)
    assert r['q'][1].sum() < r2['q'][1].sum()


def test_advance_reaches_the_interval_end_within_cfl():
    m = flowing_channel()
    o = OH.HLLCOracle(m, cfl=.9)
    r = o.advance(m.q0, m.z, 2.0)
    assert sum(r['steps']) == pytest.approx(2.0, rel=1e-13)
    assert r['retries'] >= 1 and r['rejections'][0] == 'cfl'      # the first trial is the whole interval
    assert r['cfl_max'] <= .9 * (1 + 1e-12)
    r2 = o.advance(r['q'], m.z, 2.0, 2.0)                          # the hint carries over
    assert r2['retries'] == 0


def test_product_metric_inverse_matches_oracle():
    """The product's metric (numpy pseudo-inverse of the stacked 2x2 metrics,
    as hllc.py:33-43) against the oracle's closed-form inverse."""
    from braids_hip.hllc import metric_inverse
    for m in (case_mesh(), M.lake_at_rest(), flowing_channel().renumbered()):
        a, b = metric_inverse(m), OH.metric_inverse(m)
        assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()
