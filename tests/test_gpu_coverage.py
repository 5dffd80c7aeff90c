"""Parity of the remaining variants with the FP64 oracle (limiters, viscous models, boundary
conditions, initial fields, precision), and size-independent properties at full size.

Tolerance against the oracle: 1e-12 relative L2 (FP64).  At 1M cells the oracle is too slow
for a test, so the checks there are properties that hold at any size: exact conservation on
a periodic domain, independence of the device cell order, agreement of the two kernel
families.
"""
import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
import oracle as O
from conftest import rel_l2
from helpers import vortex_case, wedge_case

pytestmark = pytest.mark.gpu
TOL = 1e-12


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def steps_both(m, p, U, n, T=1e9):
    g = F.GPUSolver2(m, p)
    g.set_state(U)
    o = O.OracleSolver(m, p)
    o.set_state(U)
    tg = to = 0.0
    for _ in range(n):
        tg, ig = g.iteration_device(tg, T)
        to, io = o.step(to, T)
    out = (tg, to, g.get_state(), o.get_state(), ig, io)
    g.close()
    return out


@pytest.mark.parametrize("gradient", [F.GRA_GREENGAUSS, F.GRA_LEASTSQUARE])
def test_venkatakrishnan_limiter(gradient):
    announce(f"Venkatakrishnan limiter, vortex 48x48, gradient={gradient}, 5 RK3 steps")
    m, p, U = vortex_case(48, gradient=gradient, limiter=F.LIM_VENKATAKRISHNAN, venkat_K=0.5)
    tg, to, Ug, Uo, _, _ = steps_both(m, p, U, 5)
    assert abs(tg / to - 1) < TOL
    assert rel_l2(Ug, Uo) < TOL


def test_navier_stokes_isothermal_wall_constant_viscosity():
    announce("NS, constant viscosity, isothermal cylinder wall (O-grid 48x24), 3 forward-Euler steps")
    m = F.Mesh.generate(F.GEN_CYLINDER_OGRID, 48, 24, a=0.5, b=8.0, c=1.12, equation=F.EQ_NS,
                        isothermal_wall=True)
    sets = [s[0] if isinstance(s, tuple) else s for s in m.boundary_sets()]
    inf = F.ruvp_from_mach(0.2, 0.0)
    Tinf = inf[3] / (inf[0] * 287.06)
    p = F.default_params(equation=F.EQ_NS, viscosity_model=F.MU_CONSTANT, mu0=inf[0] * 0.2 * 340.0 * 1.0 / 100.0,
                         inf_ruvp=inf, inlet_ruvp=inf, outlet_ruvp=inf, CFL=0.8, wall_temperature=1.2 * Tinf)
    U = F.field_init(m, p, F.INIT_UNIFORM)
    tg, to, Ug, Uo, _, _ = steps_both(m, p, U, 3)
    assert abs(tg / to - 1) < TOL
    assert rel_l2(Ug, Uo) < TOL, sets


def test_shock_tube_mesh_and_initial_field(golden_dir):
    announce("reference shocktube_100 mesh, shock-tube initial field, LLF, 10 RK3 steps")
    m = F.Mesh.read_su2(golden_dir / "shocktube_100.su2")
    p = F.default_params(flux_scheme=F.SOL_LLF, time_advance=F.EVO_RK3, CFL=0.5,
                         inlet_ruvp=[1.0, 0.0, 0.0, 1.0], outlet_ruvp=[0.125, 0.0, 0.0, 0.1],
                         inf_ruvp=[1.0, 0.0, 0.0, 1.0])
    xs = float(np.median(m.arrays()["node_xy"][0]))
    U = F.field_init(m, p, F.INIT_SHOCK_TUBE, (xs, 0.0, 1.0, 0.0))
    assert len(np.unique(U[0])) == 2
    tg, to, Ug, Uo, _, _ = steps_both(m, p, U, 10)
    assert abs(tg / to - 1) < TOL
    assert rel_l2(Ug, Uo) < TOL


def test_double_shock_reflection_boundary(tmp_path):
    announce("DSR boundary + initial field on a 64x16 channel, Roe, 5 RK3 steps")
    m0 = F.Mesh.generate(F.GEN_RECT_CHANNEL, 64, 16, a=4.0, b=1.0, jitter=0.05, seed=3)
    path = tmp_path / "dsr.su2"
    m0.write_su2(path)
    text = path.read_text().split("\n")
    tags = [k for k, l in enumerate(text) if l.startswith("MARKER_TAG")]
    text[tags[-1]] = "MARKER_TAG= dsr"
    path.write_text("\n".join(text))
    m = F.Mesh.read_su2(path)
    g30 = np.deg2rad(30.0)
    pre = [8.0, 8.25 * np.cos(g30), -8.25 * np.sin(g30), 116.5]
    post = [1.4, 0.0, 0.0, 1.0]
    p = F.default_params(flux_scheme=F.SOL_ROE, time_advance=F.EVO_RK3, CFL=0.5, inlet_ruvp=pre,
                         outlet_ruvp=post, inf_ruvp=post, dsr_shock_x=1.0 / 6.0, dsr_shock_y=0.0,
                         dsr_shock_speed=10.0)
    U = F.field_init(m, p, F.INIT_DSR, (1.0 / 6.0, 0.0, 60.0))
    tg, to, Ug, Uo, ig, io = steps_both(m, p, U, 5)
    assert ig.n_restricted == io.n_restricted
    assert abs(tg / to - 1) < TOL
    assert rel_l2(Ug, Uo) < TOL


def test_mixed_precision_vortex():
    announce("FP32-mixed vs FP64 on the periodic vortex, 10 RK3 steps")
    m, p, U = vortex_case(64)
    outs = []
    for prec in (F.PREC_FP64, F.PREC_MIXED):
        p.precision = prec
        g = F.GPUSolver2(m, p)
        g.set_state(U)
        g.advance(10, 0.0, 1e9)
        outs.append(g.get_state())
        g.close()
    for k in range(4):
        assert rel_l2(outs[1][k], outs[0][k]) < 1e-6, (k, rel_l2(outs[1][k], outs[0][k]))


def test_periodic_conservation_at_size():
    announce("conservation on a 512x512 periodic square (524,288 cells), vortex, 3 RK3 steps")
    m, p, U = vortex_case(512)
    V = m.arrays()["element_volume"]
    g = F.GPUSolver2(m, p)
    g.set_state(U)
    g.advance(3, 0.0, 1e9)
    U3 = g.get_state()
    g.close()
    before, after = U @ V, U3 @ V
    assert np.all(np.abs(after - before) <= 1e-12 * np.abs(before)), (before, after)
    assert rel_l2(U3, U) > 1e-6   # the field did move


def test_full_size_order_and_family_invariance():
    announce("wedge 1024x512 (1,048,576 cells): Hilbert vs RCM order, tile vs cell kernels, 2 RK3 steps")
    m, p, U = wedge_case(1024, 512, scheme=F.SOL_ROE, seed=2024)
    outs = {}
    for name, renumber, kernels in (("hilbert_tile", F.RENUMBER_HILBERT, F.KERNELS_TILE),
                                    ("rcm_tile", F.RENUMBER_RCM, F.KERNELS_TILE),
                                    ("hilbert_cell", F.RENUMBER_HILBERT, F.KERNELS_CELL)):
        p.renumber, p.kernels = renumber, kernels
        g = F.GPUSolver2(m, p)
        g.set_state(U)
        t, _ = g.advance(2, 0.0, 1e9)
        outs[name] = (t, g.get_state())
        g.close()
    assert outs["hilbert_tile"][0] == outs["rcm_tile"][0]
    assert np.array_equal(outs["hilbert_tile"][1], outs["rcm_tile"][1])
    assert rel_l2(outs["hilbert_cell"][1], outs["hilbert_tile"][1]) < 1e-13


TINY = {
    # name: node x, node y, element types (5 triangle, 9 quad), connectivity, markers
    "one_triangle": ([0, 1, 0], [0, 0, 1], [5], [[0, 1, 2]], [("far", [(0, 1), (1, 2), (2, 0)])]),
    "two_triangles": ([0, 1, 1, 0], [0, 0, 1, 1], [5, 5], [[0, 1, 2], [0, 2, 3]],
                      [("inlet", [(3, 0)]), ("outlet", [(1, 2)]), ("wall_slippery", [(0, 1), (2, 3)])]),
    "one_quad": ([0, 1, 1, 0], [0, 0, 1, 1], [9], [[0, 1, 2, 3]], [("far", [(0, 1), (1, 2), (2, 3), (3, 0)])]),
}


@pytest.mark.parametrize("kernels", [F.KERNELS_CELL, F.KERNELS_TILE])
@pytest.mark.parametrize("name", list(TINY))
def test_tiny_meshes(name, kernels):
    """The smallest meshes: one cell whose faces are all boundaries, two cells sharing one face,
    a quad split into two; one tile, one work-group, 3 RK3 steps against the oracle."""
    announce(f"tiny mesh {name}, kernels={kernels}, 3 RK3 steps")
    x, y, et, conn, mk = TINY[name]
    m = F.Mesh.from_lists(np.array(x, float), np.array(y, float), list(range(len(x))), et, conn, mk)
    p = F.default_params(inf_ruvp=[1.0, 0.3, 0.1, 1.0], inlet_ruvp=[1.2, 0.5, 0.0, 1.1],
                         outlet_ruvp=[1.0, 0.2, 0.0, 0.9], time_advance=F.EVO_RK3, CFL=0.5, kernels=kernels)
    U = F.field_init(m, p, F.INIT_UNIFORM)
    U[0] *= 1.1 + 0.1 * np.arange(U.shape[1])
    tg, to, Ug, Uo, _, _ = steps_both(m, p, U, 3)
    assert abs(tg / to - 1) < TOL
    assert rel_l2(Ug, Uo) < TOL
