"""Output quantities formed on the device (node fields, surface force) against the oracle,
through the C ABI.  The node field kernel uses the reference's division forms without
multiply-add contraction, so values agree to the last bit or an ulp (sqrt of Ma)."""
import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
import oracle as O
from helpers import cylinder_case

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def maxrel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def test_node_field_and_surface_force_cylinder(golden_dir):
    announce("node fields and surface force on the cylinder continue-file state")
    m, p, U = cylinder_case(golden_dir)
    g = F.GPUSolver2(m, p)
    g.set_state(U)
    R, cp, ma = g.node_field()
    Ro, cpo, mao = O.node_field(m.soa, U, p.gamma, list(p.inf_ruvp))
    assert np.array_equal(R, Ro), maxrel(R, Ro)
    assert np.array_equal(cp, cpo), maxrel(cp, cpo)
    assert maxrel(ma, mao) < 1e-15
    sets = [s[2] for s in m.boundary_sets()]
    cx, cy = g.surface_force(1.0)
    cxo, cyo = O.surface_force(m.soa, cpo, sets, 1.0)
    assert np.array_equal(cx, cxo) and np.array_equal(cy, cyo), (cx, cxo, cy, cyo)
    # after steps the node field follows the device state
    g.advance(2, 0.0, 1e9)
    U2 = g.get_state()
    R2, _, _ = g.node_field()
    R2o, _, _ = O.node_field(m.soa, U2, p.gamma, list(p.inf_ruvp))
    assert np.array_equal(R2, R2o)
    g.close()


def test_surface_force_rejects_interior_edges():
    announce("surface force: an interior edge is refused with a message")
    m = F.Mesh.generate(F.GEN_CYLINDER_OGRID, 32, 16, a=0.5, b=6.0, c=1.1)
    p = F.default_params(inf_ruvp=[1.0, 0.5, 0.0, 1.0])
    g = F.GPUSolver2(m, p)
    g.set_state(F.field_init(m, p, F.INIT_UNIFORM))
    interior = int(np.nonzero(m.arrays()["edge_setID"] == -1)[0][0])
    with pytest.raises(F.FvmrError) as e:
        g.surface_force(1.0, sets=[[interior]])
    assert "not in a boundary set" in str(e.value)
    cx, cy = g.surface_force(1.0)
    assert np.all(np.abs(cx) < 1e-12) and np.all(np.abs(cy) < 1e-12)   # uniform pressure
    g.close()


def test_checkpoint_and_restart_through_a_continue_file(golden_dir, tmp_path):
    announce("restart: reference continue file -> 2 steps -> written continue file -> 2 more steps")
    m, U, t, istep, cfl = F.Mesh.read_continue(golden_dir / "pause_cylinder_allfar_50_20000.dat", equation=F.EQ_NS)
    inf = F.ruvp_from_mach(0.5, 0.0)
    u_ref = float(np.hypot(inf[1], inf[2]))
    p = F.default_params(equation=F.EQ_NS, inf_ruvp=inf, inlet_ruvp=inf, outlet_ruvp=inf, CFL=cfl,
                         Pr=0.73, mu0=inf[0] * u_ref * 1.0 / 6.5e6)
    g = F.GPUSolver2(m, p)
    g.set_state(U)
    o = O.OracleSolver(m, p)
    o.set_state(U)
    tg = to = t
    for _ in range(2):
        tg, _ = g.iteration_device(tg, 1e9)
        to, _ = o.step(to, 1e9)
    assert abs(tg / to - 1) < 1e-12
    Ug = g.get_state()
    assert np.linalg.norm(Ug - o.get_state()) / np.linalg.norm(o.get_state()) < 1e-12
    path = tmp_path / "pause_restart[20002].dat"
    m.write_continue(path, Ug, tg, istep + 2, cfl)
    for _ in range(2):
        tg, _ = g.iteration_device(tg, 1e9)
    straight = g.get_state()
    g.close()
    m2, U2, t2, i2, _ = F.Mesh.read_continue(path, equation=F.EQ_NS)
    assert i2 == istep + 2 and np.array_equal(U2, Ug)
    g2 = F.GPUSolver2(m2, p)
    g2.set_state(U2)
    t3 = t2
    for _ in range(2):
        t3, _ = g2.iteration_device(t3, 1e9)
    assert t3 == tg
    assert np.array_equal(g2.get_state(), straight)
    g2.close()


def test_vertex_colors_match_oracle(golden_dir):
    announce("GUI vertex colours of rho, u, v, p vs the oracle, cylinder continue-file state")
    m, p, U = cylinder_case(golden_dir)
    g = F.GPUSolver2(m, p)
    g.set_state(U)
    ranges = [(0.0, 2.0), (-0.2, 0.3), (-0.1, 0.1), (0.5, 1.0)]
    cmap = F.GPUSolver2.DEFAULT_COLORMAP
    for var, (lo, hi) in enumerate(ranges):
        # primitives are formed as the reference's U2ruvp; scale the range to the state's units
        R, _, _ = O.node_field(m.soa, U, p.gamma, p.inf_ruvp[:])
        span = float(np.max(np.abs(R[var]))) or 1.0
        lo_, hi_ = lo * span, hi * span
        vert = np.zeros((m.soa.num_node, 6), np.float32)
        vert[:, :3] = 7.0
        dev = g.vertex_colors(var, lo_, hi_, cmap, vertices=vert)
        ref = O.vertex_colors(m.soa, U, p.gamma, var, lo_, hi_, cmap)
        assert np.array_equal(dev, ref), var
        assert np.all(vert[:, :3] == 7.0)                  # x, y, z untouched
    g.close()
