"""Write the golden fixtures of the river operators from the reference's own
single-source headers (river_explicit.h, river_midpoint.h) compiled with g++
by `make -C oracle ref` (oracle/_ref/libriver_*_ref.so).

Run in a container that holds /root/reference:
    python tests/golden/make_golden_river.py
Output (committed; data only): river_ops_ref.npz with, for several seeded
cases on a small triangle mesh, the operator inputs and the outputs of
explicit_frequency, midpoint_inventory, midpoint_faces and midpoint_gather.
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / 'braids-hkt_amd'))

from braids_hip import mesh as M  # noqa: E402
from oracle import river as OR  # noqa: E402

OUT = Path(__file__).resolve().parent


def case_mesh():
    return M.plain_channel(nx=12, ny=4, length=48.0, width=16.0)


def case_inputs(m, seed):
    """Random state and face data with dry cells, zero and negative water and
    a masked exterior trace; dt large enough that some cells drain."""
    rng = np.random.default_rng(seed)
    n, nf = m.nc, m.nf
    q = np.vstack([rng.uniform(0, 2, n), rng.normal(0, 1, n), rng.normal(0, 1, n)])
    q[0, rng.random(n) < 0.15] = 0.0
    q[0, rng.random(n) < 0.05] = 1e-9
    water = rng.normal(0, 2, nf)
    water[rng.random(nf) < 0.1] = 0.0
    speed = rng.uniform(0, 6, nf)
    trace = np.vstack([rng.uniform(0, 2, nf), rng.normal(0, 1, nf), rng.normal(0, 1, nf)])
    mask = ((m.neighbour < 0) & (rng.random(nf) < 0.5)).astype(np.int32)
    base = rng.uniform(0, 2, n)
    dt = float(rng.uniform(0.05, 1.0))
    rate0 = rng.normal(0, 1, (3, n))
    return dict(q=q, water=water, speed=speed, trace=trace, mask=mask, base=base, dt=dt, rate0=rate0)


def run_ops(ops, m, x, grav=9.81, dry=1e-8):
    P = OR._ptr
    n, nf = m.nc, m.nf
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    owner, nbr, bid = i32(m.owner), i32(m.neighbour), i32(m.bid)
    off, faces, signs = i32(m.offset), i32(m.faces_csr), i32(m.signs)
    area, length, normal = f64(m.area), f64(m.length), f64(m.normal)
    q = f64(x['q']); trace = f64(x['trace']); mask = i32(x['mask'])
    freq = np.zeros(n)
    ops['explicit_frequency'](n, nf, grav, dry, P(q), P(f64(x['speed'])), P(trace), P(mask), P(owner), P(normal),
                              P(area), P(length), P(off), P(faces), P(freq))
    theta = np.zeros(n)
    ops['midpoint_inventory'](n, x['dt'], P(f64(x['base'])), P(area), P(length), P(off), P(faces), P(signs),
                              P(f64(x['water'])), P(theta))
    water = f64(x['water']).copy(); corr = np.zeros(2 * nf); status = np.zeros(nf, np.int32)
    ops['midpoint_faces'](nf, n, dry, P(q), P(theta), P(owner), P(nbr), P(bid), P(water), P(corr), P(status))
    rate = f64(x['rate0']).copy()
    ops['midpoint_gather'](n, nf, P(area), P(length), P(off), P(faces), P(signs), P(water), P(corr), P(rate))
    return dict(frequency=freq, theta=theta, water_out=water, correction=corr, status=status, rate=rate)


def main():
    ops = OR.reference_ops()
    if ops is None:
        raise SystemExit('build the reference operators first: make -C oracle ref')
    m = case_mesh()
    data = {}
    for k, seed in enumerate((3, 5, 7, 11)):
        x = case_inputs(m, seed)
        y = run_ops(ops, m, x)
        for name, v in {**x, **y}.items():
            data[f'c{k}_{name}'] = np.asarray(v)
    np.savez_compressed(OUT / 'river_ops_ref.npz', **data)
    print('wrote', OUT / 'river_ops_ref.npz')


if __name__ == '__main__':
    main()
