"""Mesh contract: connectivity, incidence and renumbering are bit-exact
against the oracle's scalar restatements (oracle/mesh_oracle.py)."""
import numpy as np
import pytest

from braids_hip import mesh as M
from oracle import mesh_oracle as MO


@pytest.fixture(scope='module')
def channel():
    return M.plain_channel()


def test_plain_channel_counts(channel):
    """25,000 triangles, 38,025 faces, 1,050 boundary faces (4 m quads, 500 x 25)."""
    m = channel
    assert (m.nc, m.nf) == (25000, 38025)
    assert int((m.neighbour < 0).sum()) == 1050
    assert m.metrics()['closure_linf'] < 1e-12
    assert abs(m.metrics()['area_m2'] - 2000.0 * 100.0) < 1e-6


def test_incidence_matches_load_case_assembly(channel):
    m = channel
    off, faces, signs = MO.csr_fortran(m.nc, m.owner.tolist(), m.neighbour.tolist())
    assert np.array_equal(m.offset, off)
    assert np.array_equal(m.faces_csr, faces)
    assert np.array_equal(m.signs, signs)
    boff, bfaces = MO.boundary_fortran(m.nb, m.bid.tolist())
    assert np.array_equal(m.boundary_offset, boff)
    assert np.array_equal(m.boundary_faces, bfaces)


def test_incidence_after_renumbering_matches_load_case_assembly():
    m = M.braided_reach(nx=60, ny=12, length=300.0, width=60.0).renumbered('hilbert')
    off, faces, signs = MO.csr_fortran(m.nc, m.owner.tolist(), m.neighbour.tolist())
    assert np.array_equal(m.offset, off) and np.array_equal(m.faces_csr, faces) and np.array_equal(m.signs, signs)


def test_hilbert_keys_bit_exact():
    rng = np.random.default_rng(3)
    x = rng.uniform(-50, 5000, 3000); y = rng.uniform(0, 400, 3000)
    assert np.array_equal(M.hilbert_keys(x, y, 16), MO.hilbert_keys_scalar(x, y, 16))


def test_renumbering_is_a_consistent_permutation():
    m = M.braided_reach(nx=40, ny=10, length=200.0, width=50.0)
    r = m.renumbered('hilbert')
    assert sorted(r.cell_perm.tolist()) == list(range(m.nc))
    assert sorted(r.face_perm.tolist()) == list(range(m.nf))
    inv = np.empty(m.nc, int); inv[r.cell_perm] = np.arange(m.nc)
    # every face keeps its geometry and orientation, with renamed cells
    assert np.array_equal(r.owner, inv[m.owner[r.face_perm]])
    nb_old = m.neighbour[r.face_perm]
    assert np.array_equal(r.neighbour, np.where(nb_old >= 0, inv[np.maximum(nb_old, 0)], -1))
    assert np.array_equal(r.normal, m.normal[:, r.face_perm])
    assert np.array_equal(r.z, m.z[r.cell_perm])
    assert np.array_equal(r.to_original_cells(r.q0), m.q0)
    assert r.metrics()['closure_linf'] < 1e-12


def test_renumbered_oracle_run_maps_back():
    """Renumbering changes only the per-cell summation order: FP64 results
    agree with the original ordering to roundoff."""
    from oracle import swe
    m = M.braided_reach(nx=60, ny=12, length=300.0, width=60.0, amplitude=8.0, wavelength=150.0,
                        nbars=12, q=30.0, depth=1.0, bar_height=(0.3, 1.0), bar_sigma=(3.0, 10.0), bank=2.0)
    r = m.renumbered('hilbert')
    qa, *_ = swe.Oracle(m, scheme=5).advance(m.q0, m.z, 10 ** 9, tend=2.0)
    qb, *_ = swe.Oracle(r, scheme=5).advance(r.q0, r.z, 10 ** 9, tend=2.0)
    qb = r.to_original_cells(qb)
    for j in range(3):
        assert np.linalg.norm(qa[j] - qb[j]) <= 1e-10 * np.linalg.norm(qa[j])


def test_refined_mesh_contract():
    """A locally refined conforming triangulation: every cell is a triangle
    and closes (sum of L n over its faces is zero), normals point out of the
    owner, the areas tile the rectangle, each transition quad is a fan of
    4 + (r-1) x (refined sides) triangles."""
    nx, ny, r = 12, 6, 3
    ref = np.zeros((ny, nx), bool)
    ref[2:4, 4:7] = True
    ref[0, 0] = True
    m = M.refined_triangles(nx, ny, 2.0, 2.0, ref, r, bed=lambda x, y: 0 * x, manning=lambda x, y: 0.03 + 0 * x,
                            depth0=lambda x, y, z: 1.0 - z, inflow_q=1.0, outlet_stage=1.0)
    pad = np.pad(ref, 1)
    sides = (pad[1:-1, :-2].astype(int) + pad[1:-1, 2:] + pad[:-2, 1:-1] + pad[2:, 1:-1]) * ~ref
    fan = (sides > 0)
    assert m.nc == 2 * (~ref & ~fan).sum() + 2 * r * r * ref.sum() + (4 * fan + (r - 1) * sides).sum()
    assert np.isclose(m.area.sum(), nx * ny * 4.0, rtol=1e-14)
    vec = m.length[None, :] * m.normal
    closure = np.zeros((2, m.nc))
    inner = m.neighbour >= 0
    for a in range(2):
        closure[a] += np.bincount(m.owner, weights=vec[a], minlength=m.nc)
        closure[a] -= np.bincount(m.neighbour[inner], weights=vec[a][inner], minlength=m.nc)
    assert np.abs(closure).max() < 1e-12
    assert (((m.mid - m.xy[:, m.owner]) * m.normal).sum(0) > 0).all()
    assert (np.diff(m.offset) == 3).all()
    assert m.area.min() > 0
    assert np.isclose(m.length[m.bid == 0].sum(), ny * 2.0) and np.isclose(m.length[m.bid == 1].sum(), ny * 2.0)
    rm = m.renumbered('hilbert')
    assert np.array_equal(np.sort(rm.cell_perm), np.arange(m.nc))
