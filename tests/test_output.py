"""The oracle's output quantities (FieldWriter node fields and surface force) against an
independent numpy restatement and closed forms.  CPU only."""
import numpy as np

import cfd_solver_2d_gpu_amd as F
import oracle as O
from helpers import vortex_case


def test_node_field_matches_numpy_in_element_order():
    m, p, U = vortex_case(24)
    a = m.arrays()
    gamma = 1.4
    r = np.empty_like(U)
    r[0] = U[0]
    r[1] = U[1] / U[0]
    r[2] = U[2] / U[0]
    r[3] = r[0] * (gamma - 1) * (U[3] / U[0] - (r[1] * r[1] + r[2] * r[2]) * 0.5)
    nv = m.soa.num_node
    acc = np.zeros((4, nv))
    cnt = np.zeros(nv)
    for ie in range(m.num_element):          # element order, slot order, as the reference loops
        for j in range(3):
            v = a["element_nodes"][j, ie]
            acc[:, v] += r[:, ie]
            cnt[v] += 1
    acc /= cnt
    R, cp, ma = O.node_field(m.soa, U, gamma, [1.0, 1.0, 1.0, 1.0])
    assert np.array_equal(R, acc)
    assert np.array_equal(cp, (acc[3] - 1.0) / (0.5 * 1.0 * 2.0))
    assert np.allclose(ma, np.sqrt((acc[1] ** 2 + acc[2] ** 2) / (gamma * acc[3] / acc[0])), rtol=1e-15, atol=0)


def test_node_field_uniform_state():
    m = F.Mesh.generate(F.GEN_CYLINDER_OGRID, 32, 16, a=0.5, b=6.0, c=1.1)
    inf = [1.2, 0.3, -0.1, 0.8]
    p = F.default_params(inf_ruvp=inf)
    U = F.field_init(m, p, F.INIT_UNIFORM)
    R, cp, ma = O.node_field(m.soa, U, 1.4, inf)
    assert np.allclose(R, np.array(inf)[:, None], rtol=1e-14, atol=1e-15)
    assert np.all(np.abs(cp) < 1e-12)
    assert np.allclose(ma, np.hypot(0.3, -0.1) / np.sqrt(1.4 * 0.8 / 1.2), rtol=1e-13)


def test_surface_force_closed_forms():
    m = F.Mesh.generate(F.GEN_CYLINDER_OGRID, 48, 16, a=0.5, b=6.0, c=1.1)
    a = m.arrays()
    x, y = a["node_xy"]
    sets = [s[2] for s in m.boundary_sets()]
    names = [s[0] for s in m.boundary_sets()]
    wall = sets[names.index("wall")]
    # Cp = x (or y) at the nodes: the trapezoid rule of x n_x dS is exact, the shoelace area
    n0, n1 = a["edge_nodes"][0][wall], a["edge_nodes"][1][wall]
    area = np.sum(0.5 * (x[n0] + x[n1]) * (y[n1] - y[n0]))
    assert abs(abs(area) - 24 * 0.25 * np.sin(2 * np.pi / 48)) < 1e-12   # regular 48-gon, r = 0.5
    cd, _ = O.surface_force(m.soa, x, [wall], 2.0)
    _, cl = O.surface_force(m.soa, y, [wall], 2.0)
    assert abs(cd[0] - area / 2.0) < 1e-13 and abs(cl[0] - area / 2.0) < 1e-13
    # uniform Cp on a closed boundary: no force
    cd, cl = O.surface_force(m.soa, np.full(len(x), 0.7), sets, 1.0)
    assert np.all(np.abs(cd) < 1e-13) and np.all(np.abs(cl) < 1e-13)


CMAP = F.GPUSolver2.DEFAULT_COLORMAP      # CudaGrid.cpp:33-39: red, yellow, green, blue, magenta


def test_vertex_colors_closed_forms():
    """The oracle's GPU::setVerticesColor: control points give their colours, midpoints of an
    interval the mean of its ends, out-of-range values clamp to the end colours, inputs outside
    the control points purple, and a node takes the largest value of its cells."""
    m = F.Mesh.generate(F.GEN_CYLINDER_OGRID, 24, 8, a=0.5, b=4.0, c=1.1)
    nv, ne = m.soa.num_node, m.num_element
    a = m.arrays()
    cm = np.array(CMAP, dtype=np.float32)

    def colors(rho, lo=0.0, hi=1.0, cmap=cm):
        U = np.stack([np.full(ne, rho) if np.isscalar(rho) else rho, np.zeros(ne), np.zeros(ne), np.full(ne, 2.5)])
        return O.vertex_colors(m.soa, U, 1.4, 0, lo, hi, cmap)

    for k, (r, g, b, w) in enumerate(CMAP):       # each control point, up to the rounding of rho n (1/n)
        assert np.allclose(colors(w if w > 0 else 0.0), [r, g, b], atol=2e-6), k
    mid = colors(0.375)                              # halfway from yellow to green
    assert np.allclose(mid, [0.5, 1.0, 0.0], atol=1e-6)
    assert np.array_equal(colors(7.0), np.tile(cm[-1, :3], (nv, 1)))     # clamped high: magenta
    assert np.array_equal(colors(0.5, lo=2.0, hi=3.0), np.tile(cm[0, :3], (nv, 1)))   # clamped low: red
    assert np.array_equal(colors(0.0, lo=0.0, hi=0.0), np.tile(cm[0, :3], (nv, 1)))   # empty range -> 0
    gap = np.array([[1, 0, 0, 0.2], [0, 0, 1, 0.8]], dtype=np.float32)
    purple = np.array([252, 3, 253], dtype=np.float32) / np.float32(255)
    assert np.array_equal(colors(0.1, cmap=gap), np.tile(purple, (nv, 1)))
    # one cell at 1 over 0.5 elsewhere: its three nodes magenta, the rest green
    rho = np.full(ne, 0.5)
    rho[5] = 1.0
    c = colors(rho)
    hot = set(a["element_nodes"][:3, 5].tolist())
    for v in range(nv):
        assert np.allclose(c[v], [1, 0, 1] if v in hot else [0, 1, 0], atol=2e-6), v
