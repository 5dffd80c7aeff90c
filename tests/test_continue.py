"""Continue files (pause_*.dat): reading the one the reference wrote, writing in its layout,
and bit-exact round trips.  CPU only."""
import numpy as np

import cfd_solver_2d_gpu_amd as F
from helpers import read_pause, vortex_case

TOPO = ("element_nodes", "element_edges", "element_neighbors", "edge_nodes", "edge_setID",
        "edge_periodicPair", "edge_elementL", "edge_elementR", "boundary_type")


def same_topology(a, b):
    for k in TOPO:
        assert np.array_equal(a[k], b[k]), k


def test_reads_the_reference_continue_file(golden_dir):
    path = golden_dir / "pause_cylinder_allfar_50_20000.dat"
    m, U, t, istep, cfl = F.Mesh.read_continue(path, equation=F.EQ_NS)
    assert (t, istep, cfl) == (0.120747, 20000, 2.2)
    (t0, i0, c0), nodes, elems, sets = read_pause(path)
    assert np.array_equal(U, elems[:, 4:8].T)
    assert np.array_equal(m.element_node_ids(), elems[:, 1:4].astype(np.int32))
    # same topology as the SU2 mesh the reference started from (coordinates in the pause file
    # carry 6 significant digits only)
    su2 = F.Mesh.read_su2(golden_dir / "cylinder_allfar_50.su2", equation=F.EQ_NS)
    same_topology(m.arrays(), su2.arrays())
    got = [(n, t, e.tolist()) for n, t, e in m.boundary_sets()]
    ref = [(n, t, e.tolist()) for n, t, e in su2.boundary_sets()]
    assert got == ref
    assert [s[0] for s in sets] == [g[0] for g in got]
    assert [s[1] for s in sets] == [[e + 1 for e in g[2]] for g in got]


def test_written_layout_matches_the_reference(golden_dir, tmp_path):
    path = golden_dir / "pause_cylinder_allfar_50_20000.dat"
    m, U, t, istep, cfl = F.Mesh.read_continue(path, equation=F.EQ_NS)
    out = tmp_path / "pause_out.dat"
    m.write_continue(out, U, t, istep, cfl)
    ref = path.read_text().split("\n")
    new = out.read_text().split("\n")
    assert len(ref) == len(new)
    b_ref, b_new = ref.index("boundaries: ID, name; edgeIDs"), new.index("boundaries: ID, name; edgeIDs")
    assert b_ref == b_new and ref[b_ref:] == new[b_new:]       # sets and edge IDs, byte for byte
    for k in (0, 2, 3592):                                     # section headers
        assert ref[k] == new[k]
    for r, n in zip(ref[3:3592] + ref[3593:b_ref], new[3:3592] + new[3593:b_new]):
        rw, nw = r.split(), n.split()
        assert rw[0] == nw[0] and len(rw) == len(nw)           # IDs and columns
        assert all(abs(float(a) - float(b)) <= 5e-6 * max(1.0, abs(float(a))) for a, b in zip(rw[1:], nw[1:]))


def test_round_trip_is_bit_exact(tmp_path):
    m, p, U = vortex_case(32)
    U = U * (1 + 1e-3 * np.sin(np.arange(U.shape[1])))[None, :]
    path = tmp_path / "pause_vortex[7].dat"
    m.write_continue(path, U, 0.123456789012345678, 7, 0.5)
    m2, U2, t, istep, cfl = F.Mesh.read_continue(path)
    assert (t, istep, cfl) == (0.123456789012345678, 7, 0.5)
    assert np.array_equal(U2, U)
    a, b = m.arrays(), m2.arrays()
    same_topology(a, b)
    for k in ("node_xy", "element_xy", "element_volume", "edge_length", "edge_xy", "edge_normal"):
        assert np.array_equal(a[k], b[k]), k
    assert [(n, t_) for n, t_, _ in m.boundary_sets()] == [(n, t_) for n, t_, _ in m2.boundary_sets()]
