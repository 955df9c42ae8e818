"""Write the golden fixtures of the HLLC operators from the reference build
(oracle/_ref: g++ builds of hllc.h, hllc_turbines.h and hllc_advanced.h, made
by `make -C oracle ref`).

    python tests/golden/make_golden_hllc.py
Output (committed; data only): hllc_ref.npz.  Each case stores the mesh
geometry, the inputs and the outputs of one chained pass
reconstruct -> predict_source -> faces -> turbine_pressure -> budget ->
limit -> update (plus advanced_frequency) as the reference computes them.
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / 'braids-hkt_amd'))

from braids_hip import mesh as M  # noqa: E402
from oracle import hllc as OH  # noqa: E402

OUT = Path(__file__).resolve().parent
CASES = ((11, None, None), (12, None, None), (13, None, None), (14, 1.2, None), (15, 0.1, 7))
GEOM = ('xy', 'mid', 'normal', 'owner', 'neighbour', 'length', 'area', 'off', 'faces', 'signs', 'inverse',
        'selected')


def case_mesh(nx=10, ny=3):
    """A 40 m x 12 m channel (inlet, outlet, two walls) with a bed mound."""
    m = M.plain_channel(nx=nx, ny=ny, length=40.0, width=12.0, depth=1.0, q=6.0)
    x, y = m.xy
    m.z[:] = m.z + 0.8 * np.exp(-((x - 20.0) ** 2 + (y - 6.0) ** 2) / 20.0)
    return m


def chain_inputs(m, seed, dt=None, nan_cell=None):
    """A state with wet, dry and nearly dry cells, dry banks above the water
    line, fast cells (supercritical faces) and an empty cell next to a
    prescribed outflow (unfunded exterior withdrawal)."""
    rng = np.random.default_rng(seed)
    n, nf = m.nc, m.nf
    z = m.z.copy()
    eta = 1.0 + 0.05 * rng.standard_normal(n)
    h = np.maximum(eta - z, 0.0)
    h[rng.random(n) < 0.12] = 0.0
    h[rng.random(n) < 0.05] = 5e-9
    u = rng.normal(0.6, 0.8, n); v = rng.normal(0.0, 0.4, n)
    u[rng.random(n) < 0.1] *= 8.0
    q = np.vstack([h, h * u, h * v])
    sel = OH.open_faces(m)
    ext = np.zeros((3, nf)); tr = np.zeros((3, nf))
    fs = np.flatnonzero(sel)
    ext[:, fs] = rng.normal(0, 1.0, (3, fs.size))
    tr[0, fs] = rng.uniform(0, 1.2, fs.size); tr[1:, fs] = rng.normal(0, 1.0, (2, fs.size))
    tr[0, fs[:2]] = 0.0
    # an empty owner of a prescribed outflow face
    o = m.owner[fs[-1]]
    q[:, o] = 0.0
    ext[0, fs[-1]] = abs(ext[0, fs[-1]]) + 0.5
    rate = np.zeros((3, n))
    k = rng.random(n) < 0.3
    rate[1:, k] = rng.normal(0, 0.05, (2, int(k.sum())))
    if nan_cell is not None:
        q[1, nan_cell] = np.nan                    # a non-finite predictor
    dt = float(rng.uniform(0.02, 0.2)) if dt is None else dt
    return dict(q=q, z=z, dt=dt, rate=rate, external=ext, trace=tr,
                metric=rng.uniform(0.1, 2.0, nf), nu=float(rng.uniform(0.0, 0.5)), grav=9.81, dry=1e-8)


def run_chain(ops, g, x, pressure=True):
    n = g['area'].size; nf = g['length'].size
    c = lambda a, t=np.float64: np.ascontiguousarray(a, dtype=t)
    q, z = c(x['q']), c(x['z'])
    grav, dry, dt = float(x['grav']), float(x['dry']), float(x['dt'])
    grad = np.zeros((8, n)); half = np.zeros((4, n)); qhalf = np.zeros((3, n)); pstat = np.zeros(n, np.int32)
    ops['hllc_reconstruct'](n, nf, grav, dry, dt, q, z, g['xy'], g['mid'], g['normal'], g['owner'],
                            g['neighbour'], g['selected'], g['off'], g['faces'], g['inverse'], grad, half, qhalf,
                            pstat)
    out = dict(grad=grad.copy(), half0=half.copy(), qhalf0=qhalf.copy(), pstatus=pstat.copy())
    ops['hllc_predict_source'](n, dry, dt, c(x['rate']), half, qhalf)
    out.update(half=half.copy(), qhalf=qhalf.copy())
    fl = np.zeros((3, nf)); fr = np.zeros((3, nf)); speed = np.zeros(nf)
    ops['hllc_faces'](n, nf, grav, dry, half, grad, g['xy'], g['mid'], g['normal'], g['owner'], g['neighbour'],
                      g['selected'], c(x['external']), c(x['trace']), fl, fr, speed)
    out.update(fl0=fl.copy(), fr0=fr.copy(), speed=speed.copy())
    if pressure:
        ops['hllc_turbine_pressure'](n, nf, grav, dry, half, grad, g['xy'], g['mid'], g['normal'], g['owner'],
                                     g['neighbour'], fl, fr)
    out.update(fl1=fl.copy(), fr1=fr.copy())
    theta = np.zeros(n); freq = np.zeros(n)
    ops['hllc_budget'](n, nf, dt, q, g['area'], g['length'], g['off'], g['faces'], g['signs'], fl, speed,
                       g['selected'], pstat, theta, freq)
    out.update(theta=theta.copy(), frequency=freq.copy())
    fstat = np.zeros(nf, np.int32)
    ops['hllc_limit'](n, nf, dry, q, theta, g['owner'], g['neighbour'], g['selected'], fl, fr, fstat)
    out.update(fl=fl.copy(), fr=fr.copy(), fstatus=fstat.copy())
    nxt = np.zeros((3, n)); stat = np.zeros(n)
    ops['hllc_update'](n, nf, dt, dry, q, g['area'], g['length'], g['off'], g['faces'], g['signs'], fl, fr, fstat,
                       nxt, stat)
    out.update(next=nxt, status=stat)
    adv = np.zeros(n)
    ops['hllc_advanced_frequency'](n, nf, dry, float(x['nu']), q, c(x['metric']), g['owner'], g['neighbour'],
                                   g['off'], g['faces'], g['length'], g['area'], adv)
    out.update(advanced=adv)
    return out


def main():
    ops = OH.reference_ops()
    if ops is None:
        raise SystemExit('build the reference first: make -C oracle ref')
    m = case_mesh()
    g = OH.geometry(m)
    data = {f'g_{k}': g[k] for k in GEOM}
    for k, (seed, dt, nan_cell) in enumerate(CASES):
        x = chain_inputs(m, seed, dt, nan_cell)
        y = run_chain(ops, g, x)
        for name, val in x.items():
            data[f'c{k}_in_{name}'] = np.asarray(val)
        for name, val in y.items():
            data[f'c{k}_out_{name}'] = np.asarray(val)
    np.savez_compressed(OUT / 'hllc_ref.npz', **data)
    print('wrote', OUT / 'hllc_ref.npz')


if __name__ == '__main__':
    main()
