"""Write the golden fixtures of the Exner pieces from the reference builds
(oracle/_ref: the amdflang build of braids_bedload_mpm_module.f90 and the g++
build of bend_methods_v1_kernels.h, made by `make -C oracle ref`).

    python tests/golden/make_golden_exner.py
Output (committed; data only): exner_ref.npz with MPM inputs/outputs and the
divergence / donor_factors / limit_flux operator cases.
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / 'braids-hkt_amd'))

from braids_hip import mesh as M  # noqa: E402
from oracle import exner as OX  # noqa: E402

OUT = Path(__file__).resolve().parent


def mpm_cases(n=2000, seed=21):
    rng = np.random.default_rng(seed)
    u = rng.normal(0, 2, n); v = rng.normal(0, 1, n)
    cf = rng.uniform(0, 0.02, n); d50 = rng.uniform(1e-4, 0.05, n)
    u[:20] = 0.0; v[:20] = 0.0                      # still water
    cf[20:40] = 0.0                                 # no friction
    u[40:60] = 1e-3                                 # below threshold
    return np.vstack([u, v, cf, d50]).T


def case_mesh():
    return M.plain_channel(nx=12, ny=4, length=48.0, width=16.0)


def op_inputs(m, seed):
    rng = np.random.default_rng(seed)
    flux = rng.normal(0, 1e-3, (3, m.nf)); flux[:, rng.random(m.nf) < 0.1] = 0.0
    available = rng.uniform(-1e-3, 5e-3, m.nc)
    return dict(flux=flux, available=available, dt=float(rng.uniform(0.1, 5.0)))


def run_ops(ops, m, x, component=2):
    P = OX._ptr
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    n, nf = m.nc, m.nf
    flux = np.ascontiguousarray(x['flux'], dtype=np.float64).copy()
    fac = np.zeros(n); div = np.zeros(n)
    off, faces, signs = i32(m.offset), i32(m.faces_csr), i32(m.signs)
    ops['donor_factors'](n, nf, component, x['dt'], P(np.ascontiguousarray(x['available'])), P(flux), P(m.length),
                         P(m.area), P(off), P(faces), P(signs), P(fac))
    ops['limit_flux'](nf, component, P(i32(m.owner)), P(i32(m.neighbour)), P(fac), P(flux))
    ops['divergence'](n, nf, component, P(flux), P(m.length), P(m.area), P(off), P(faces), P(signs), P(div))
    return dict(factors=fac, limited=flux, divergence=div)


def main():
    ref = OX.reference_mpm(); ops = OX.bed_ops(reference=True)
    if ref is None or ops is None:
        raise SystemExit('build the reference first: make -C oracle ref')
    X = mpm_cases()
    Y = np.array([OX.mpm(u, v, cf, d50, fn=ref) for u, v, cf, d50 in X])
    data = dict(mpm_in=X, mpm_out=Y)
    m = case_mesh()
    for k, seed in enumerate((1, 2, 3)):
        x = op_inputs(m, seed)
        y = run_ops(ops, m, x)
        for name, val in {**x, **y}.items():
            data[f'c{k}_{name}'] = np.asarray(val)
    np.savez_compressed(OUT / 'exner_ref.npz', **data)
    print('wrote', OUT / 'exner_ref.npz')


if __name__ == '__main__':
    main()
