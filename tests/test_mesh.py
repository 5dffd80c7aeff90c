"""Mesh preprocessing: the hashed builder of the library against the oracle's restatement
of the reference's linear-search builder (bit-exact), against the reference-written
continue file, and size-independent invariants at the 1M-cell workload."""
import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
import oracle as O
from helpers import read_pause


def same_arrays(a, b):
    bad = []
    for k in a:
        x, y = a[k], b[k]
        if x.shape != y.shape or x.tobytes() != y.tobytes():
            bad.append(k)
    return bad


@pytest.mark.parametrize("name,eq", [("ivortex50.su2", F.EQ_EULER), ("cylinder_allfar_50.su2", F.EQ_NS),
                                     ("shocktube_100.su2", F.EQ_EULER), ("neuvillette.su2", F.EQ_NS),
                                     ("dsr131new_slippery.su2", F.EQ_EULER)])
def test_builder_bit_exact_on_reference_meshes(golden_dir, name, eq):
    m = F.Mesh.read_su2(golden_dir / name, equation=eq)
    o = O.OracleMesh.read_su2(golden_dir / name, equation=eq)
    assert same_arrays(m.arrays(), o.arrays()) == []
    assert np.array_equal(m.element_node_ids(), o.element_node_ids())
    mb, ob = m.boundary_sets(), o.boundary_sets()
    assert [(n, t) for n, t, _ in mb] == [(n, t) for n, t, _ in ob]
    assert all(np.array_equal(a[2], b[2]) for a, b in zip(mb, ob))


def test_builder_matches_reference_continue_file(golden_dir):
    """pause_cylinder_allfar_50[20000].dat was written by the reference: its element node lists
    are post-CCW-swap and its boundary sections list reference edge IDs (1-based)."""
    (t, istep, cfl), nodes, elems, sets = read_pause(golden_dir / "pause_cylinder_allfar_50_20000.dat")
    assert istep == 20000
    m = F.Mesh.read_su2(golden_dir / "cylinder_allfar_50.su2", equation=F.EQ_NS)
    assert np.array_equal(m.element_node_ids(), elems[:, 1:4].astype(int))
    bs = m.boundary_sets()
    assert [b[0] for b in bs] == [s[0] for s in sets] == ["farfield", "wall"]
    for (name, typ, edges), (_, ids) in zip(bs, sets):
        assert np.array_equal(edges + 1, np.array(ids))
    a = m.arrays()
    # coordinates are written with 6 significant digits
    assert np.allclose(a["node_xy"].T, nodes[:, 1:3], rtol=2e-5, atol=1e-5)
    assert [t for _, t, _ in bs] == [F.BC_INF, F.BC_WALL_ADIABAT]


@pytest.mark.parametrize("family,ni,nj,kw", [
    (F.GEN_PERIODIC_SQUARE, 64, 64, dict(a=0.0, b=10.0)),
    (F.GEN_WEDGE_CHANNEL, 48, 24, dict(a=3.0, b=1.0, c=0.5, d=15.0, jitter=0.1, seed=3)),
    (F.GEN_CYLINDER_OGRID, 48, 24, dict(a=0.5, b=30.0, c=1.05)),
    (F.GEN_NACA_OGRID, 64, 24, dict(b=20.0, c=1.1, naca_digits=12)),
])
def test_generated_meshes_match_oracle_through_su2(tmp_path, family, ni, nj, kw):
    m = F.Mesh.generate(family, ni, nj, **kw)
    path = tmp_path / "m.su2"
    m.write_su2(path)
    again = F.Mesh.read_su2(path)
    o = O.OracleMesh.read_su2(path)
    assert same_arrays(m.arrays(), again.arrays()) == []
    assert same_arrays(m.arrays(), o.arrays()) == []


def check_invariants(m):
    a = m.arrays()
    nc, nf, nv = m.num_element, m.num_edge, m.num_node
    L, R = a["edge_elementL"], a["edge_elementR"]
    faces = a["element_edges"][:3]
    # every edge is listed by its owners: interior twice, boundary (incl. periodic) once
    cnt = np.bincount(faces.ravel(), minlength=nf)
    per = a["edge_periodicPair"] >= 0
    interior = (R >= 0) & ~per
    assert np.all(cnt[interior] == 2) and np.all(cnt[~interior] == 1)
    assert np.all(a["element_edges"][3] == -1) and np.all(a["element_nodes"][3] == -1)
    # closure: sum over faces of s * n * length vanishes per cell
    c = np.arange(nc)
    sx = np.zeros(nc)
    sy = np.zeros(nc)
    for j in range(3):
        e = faces[j]
        s = np.where(R[e] != c, 1.0, -1.0)
        sx += s * a["edge_normal"][0][e] * a["edge_length"][e]
        sy += s * a["edge_normal"][1][e] * a["edge_length"][e]
    h = np.sqrt(a["element_volume"])
    assert np.max(np.abs(sx) / h) < 1e-6 and np.max(np.abs(sy) / h) < 1e-6  # lengths carry float rounding
    # normals are unit up to the float rounding of the length, point from L to R
    nn = np.hypot(*a["edge_normal"])
    assert np.max(np.abs(nn - 1)) < 1e-7
    assert np.all(a["element_volume"] > 0)
    # no overlapping cells: the areas sum to the domain's area, the boundary integral of x n_x
    # (boundary edges have the domain outside their R side)
    b = R < 0
    b |= per
    area = np.sum(a["edge_xy"][0][b] * a["edge_normal"][0][b] * a["edge_length"][b])
    assert abs(a["element_volume"].sum() / area - 1) < 1e-6
    # periodic pairs are symmetric translates and R is the pair's L
    pe = a["edge_periodicPair"][per]
    assert np.array_equal(a["edge_periodicPair"][pe], np.nonzero(per)[0])
    assert np.array_equal(R[per], L[pe])
    # Euler characteristic of a disk (an O-grid is an annulus: 0)
    return nv - nf + nc


def test_invariants_on_the_1M_wedge():
    m = F.Mesh.generate(F.GEN_WEDGE_CHANNEL, 1024, 512, a=3.0, b=1.0, c=0.5, d=15.0, jitter=0.1, seed=2024)
    assert (m.num_element, m.num_edge, m.num_node) == (1048576, 1574400, 525825)
    assert check_invariants(m) == 1


def test_invariants_on_periodic_and_ogrid():
    assert check_invariants(F.Mesh.generate(F.GEN_PERIODIC_SQUARE, 64, 64, b=10.0)) == 1
    assert check_invariants(F.Mesh.generate(F.GEN_CYLINDER_OGRID, 64, 32, a=0.5, b=30.0, c=1.02)) == 0
    assert check_invariants(F.Mesh.generate(F.GEN_NACA_OGRID, 64, 24, b=20.0, c=1.1, naca_digits=12)) == 0
    assert check_invariants(F.Mesh.generate(F.GEN_NACA_OGRID, 64, 24, b=20.0, c=1.1, naca_digits=2412)) == 0


def test_builder_errors(tmp_path):
    x = np.array([0, 1, 1, 0.0])
    y = np.array([0, 0, 1, 1.0])
    ids = np.arange(4)
    et = [9]
    ec = [[0, 1, 2, 3]]
    with pytest.raises(F.FvmrError, match="not paired|pairs"):
        F.Mesh.from_lists(x, y, ids, et, ec, [("periodic_0", [(0, 1)])])
    with pytest.raises(F.FvmrError, match="unknown boundary type"):
        F.Mesh.from_lists(x, y, ids, et, ec, [("nonsense", [(0, 1)])])
    with pytest.raises(F.FvmrError, match="pEdge == nullptr"):
        F.Mesh.from_lists(x, y, ids, et, ec, [("wall", [(0, 2), (2, 0)])])
    with pytest.raises(F.FvmrError, match="counts differ"):
        F.Mesh.from_lists(x, y, ids, et, ec, [("periodic_0", [(0, 1)]), ("periodic_0", [(2, 3), (3, 0)])])
    # the oracle reports the same failures
    with pytest.raises(RuntimeError):
        O.OracleMesh.from_lists(x, y, ids, et, ec, [("periodic_0", [(0, 1)])])


def test_ccw_swap_and_quad_split():
    # a clockwise triangle is swapped (nodes 1 <-> 2); a quad is split (0,1,3),(3,1,2)
    x = np.array([0, 1, 0, 1.0, 2.0])
    y = np.array([0, 0, 1, 1.0, 0.0])
    m = F.Mesh.from_lists(x, y, np.arange(5), [5, 9], [[0, 2, 1, -1], [1, 4, 3, 3]], [])
    ids = m.element_node_ids()
    assert ids.tolist()[0] == [1, 2, 3]
    o = O.OracleMesh.from_lists(x, y, np.arange(5), [5, 9], [[0, 2, 1, -1], [1, 4, 3, 3]], [])
    assert np.array_equal(ids, o.element_node_ids())
