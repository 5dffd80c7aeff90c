"""Pins the CPU oracle with closed forms, invariants and textbook special cases, each
chosen so that a plausible slip (dropped term, wrong sign or index, transposed operand)
fails.  The reference ships no tests or golden vectors for this path; the one number its own
files give on its own inputs (a logged step-1 residual of the shock tube) is checked last, and
the reference-held data files are used in test_mesh.py and the GPU parity tests."""
import os

import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
import oracle as O
from helpers import vortex_case

G = 1.4


def phys_flux(U, nx, ny, g=G):
    rho, ru, rv, rE = U
    u, v = ru / rho, rv / rho
    p = (g - 1) * (rE - 0.5 * rho * (u * u + v * v))
    un = u * nx + v * ny
    return np.array([rho * un, ru * un + p * nx, rv * un + p * ny, (rE + p) * un])


def state(rho, u, v, p, g=G):
    return np.array([rho, rho * u, rho * v, p / (g - 1) + 0.5 * rho * (u * u + v * v)])


STATES = [state(1.0, 0.3, -0.2, 1.0), state(1.4, 2.0, 0.1, 1.0), state(0.5, -0.7, 0.4, 0.3)]
NORMALS = [(1.0, 0.0), (0.6, 0.8), (-np.sqrt(0.5), np.sqrt(0.5))]


def test_conversions_round_trip():
    for U in STATES:
        assert np.allclose(O.ruvp2u(O.u2ruvp(U)), U, rtol=1e-15, atol=1e-15)
    r = O.u2ruvp(state(1.2, 0.5, -0.25, 0.9))
    assert np.allclose(r, [1.2, 0.5, -0.25, 0.9], rtol=1e-14)


@pytest.mark.parametrize("solver", [O.roe, O.llf])
def test_riemann_consistency_and_conservation(solver):
    for U in STATES:
        for nx, ny in NORMALS:
            # consistency: F(U, U) = length * physical flux
            assert np.allclose(solver(U, U, nx, ny, 0.7), 0.7 * phys_flux(U, nx, ny), rtol=1e-13, atol=1e-13)
    for UL, UR in [(STATES[0], STATES[1]), (STATES[1], STATES[2]), (STATES[2], STATES[0])]:
        for nx, ny in NORMALS:
            # conservation: the flux seen from the other side with the flipped normal is opposite
            f = solver(UL, UR, nx, ny, 1.3)
            g = solver(UR, UL, -nx, -ny, 1.3)
            assert np.allclose(f, -g, rtol=1e-12, atol=1e-12)


def test_roe_is_upwind_for_supersonic_mass_and_momentum():
    # all eigenvalues positive: the dissipation makes mass and momentum the left flux;
    # the energy component keeps the reference's own energy jump (documented quirk)
    UL = state(1.0, 3.0, 0.1, 1.0)
    UR = state(1.1, 2.9, 0.2, 1.2)
    f = O.roe(UL, UR, 1.0, 0.0, 1.0)
    assert np.allclose(f[:3], phys_flux(UL, 1.0, 0.0)[:3], rtol=1e-12)


def test_llf_closed_form():
    UL, UR = STATES[0], STATES[2]
    nx, ny = 0.6, 0.8
    rl, rr = O.u2ruvp(UL), O.u2ruvp(UR)
    lam = max(abs(rl[1] * nx + rl[2] * ny) + np.sqrt(G * rl[3] / rl[0]),
              abs(rr[1] * nx + rr[2] * ny) + np.sqrt(G * rr[3] / rr[0]))
    expect = 0.5 * (phys_flux(UL, nx, ny) + phys_flux(UR, nx, ny) - lam * (UR - UL)) * 2.0
    assert np.allclose(O.llf(UL, UR, nx, ny, 2.0), expect, rtol=1e-13)


@pytest.mark.parametrize("nx,ny", [(1.0, 0.0), (-1.0, 0.0), (0.6, 0.8), (-0.8, 0.6)])
@pytest.mark.parametrize("mach", [0.3, 2.0])
def test_farfield_ghost_of_freestream_is_freestream(nx, ny, mach):
    # all four regimes (sub/supersonic, inflow/outflow) return the freestream itself
    inf = np.array([1.0, mach * np.sqrt(G), 0.0, 1.0])
    U = O.ruvp2u(inf)
    assert np.allclose(O.farfield(U, inf, nx, ny), U, rtol=1e-13, atol=1e-14)


def test_slip_wall_reflects_normal_velocity():
    U = state(1.3, 0.4, -0.9, 0.8)
    nx, ny = 0.6, 0.8
    g = O.slip(U, nx, ny)
    r, rg = O.u2ruvp(U), O.u2ruvp(g)
    un, ut = r[1] * nx + r[2] * ny, r[1] * ny - r[2] * nx
    assert np.isclose(rg[1] * nx + rg[2] * ny, -un) and np.isclose(rg[1] * ny - rg[2] * nx, ut)
    assert g[0] == U[0] and g[3] == U[3]


def test_inv2x2_correct_and_reference_bug():
    A = np.array([[2.0, 1.0], [1.0, 3.0]])
    good = O.inv2x2(A).reshape(2, 2)
    assert np.allclose(good @ A, np.eye(2), atol=1e-14)
    bug = O.inv2x2(A, compat=True).reshape(2, 2)
    # MatrixKernel.h:48-49: every det >= EPSILON is replaced by -EPSILON
    assert np.allclose(bug, -np.array([[3.0, -1.0], [-1.0, 2.0]]) / 1e-10)
    small = O.inv2x2(np.array([[1e-6, 0], [0, 1e-6]])).reshape(2, 2)
    assert np.allclose(small, np.eye(2) / 1e-10 * 1e-6)


def linear_field(m, a=(1.0, 0.3, -0.2), b=(0.5, 0.1, 0.05), c=(0.2, -0.1, 0.3), d=(2.5, 0.2, 0.1)):
    xy = m.arrays()["element_xy"]
    U = np.stack([k[0] + k[1] * xy[0] + k[2] * xy[1] for k in (a, b, c, d)])
    return U, np.array([[k[1] for k in (a, b, c, d)], [k[2] for k in (a, b, c, d)]])


def test_gradients_constant_and_linear_fields():
    m = F.Mesh.generate(F.GEN_WEDGE_CHANNEL, 24, 12, a=3.0, b=1.0, c=0.5, d=15.0, jitter=0.1, seed=5)
    a = m.arrays()
    p = F.default_params(limiter=F.LIM_NONE)
    s = O.OracleSolver(m, p)
    s.set_state(np.ones((4, m.num_element)) * np.array([[1.0], [0.2], [0.1], [2.5]]))
    gx, gy = s.gradient()
    assert np.abs(gx).max() < 1e-12 and np.abs(gy).max() < 1e-12   # GG of a constant is zero everywhere
    U, grad = linear_field(m)
    nvalid = (a["element_neighbors"][:3] >= 0).sum(axis=0)
    for gtype in (F.GRA_LEASTSQUARE,):
        p = F.default_params(limiter=F.LIM_NONE, gradient=gtype)
        s = O.OracleSolver(m, p)
        s.set_state(U)
        gx, gy = s.gradient()
        ok = nvalid >= 2   # LSQ is exact for linear fields with two independent neighbours
        assert np.allclose(gx[:, ok], grad[0][:, None], atol=1e-10)
        assert np.allclose(gy[:, ok], grad[1][:, None], atol=1e-10)


def test_barth_bounds_vertex_values():
    m, p, U = vortex_case(32, limiter=F.LIM_BARTH)
    a = m.arrays()
    s = O.OracleSolver(m, p)
    s.set_state(U)
    gx, gy = s.gradient()
    p0 = F.default_params(inf_ruvp=[1, 1, 1, 1], limiter=F.LIM_NONE)
    s0 = O.OracleSolver(m, p0)
    s0.set_state(U)
    rx, ry = s0.gradient()
    phi = np.where(np.abs(rx) > 0, gx / np.where(rx == 0, 1, rx), 1.0)
    assert np.all(phi <= 1 + 1e-12) and np.all(phi >= -1e-12)
    nb = a["element_neighbors"][:3]
    xy = a["element_xy"]
    for k in range(4):
        vals = np.concatenate([U[k][None], np.where(nb >= 0, U[k][nb], U[k][None])])
        lo, hi = vals.min(axis=0), vals.max(axis=0)
        for j in range(3):
            nd = a["element_nodes"][j]
            v = U[k] + gx[k] * (a["node_xy"][0][nd] - xy[0]) + gy[k] * (a["node_xy"][1][nd] - xy[1])
            assert np.all(v <= hi + 1e-9) and np.all(v >= lo - 1e-9)


def test_venkatakrishnan_does_not_amplify():
    m, p, U = vortex_case(32, limiter=F.LIM_VENKATAKRISHNAN)
    s = O.OracleSolver(m, p)
    s.set_state(U)
    gx, _ = s.gradient()
    s0 = O.OracleSolver(m, F.default_params(inf_ruvp=[1, 1, 1, 1], limiter=F.LIM_NONE))
    s0.set_state(U)
    rx, _ = s0.gradient()
    assert np.all(np.abs(gx) <= np.abs(rx) * (1 + 1e-12) + 1e-300)


def test_dt_closed_form_on_periodic_square_at_rest():
    """u = v = 0: alpha_e = c sqrt(L).  A corner cell has two periodic faces, each counted a
    second time through its pair edge (CalculateDt.cpp:51-52 with R = pair's L), which makes
    it the minimum: dt = CFL (h^2/2) / (c (4 sqrt(h) + sqrt(sqrt(2) h)))."""
    n, Lbox = 16, 10.0
    m = F.Mesh.generate(F.GEN_PERIODIC_SQUARE, n, n, b=Lbox)
    p = F.default_params(inf_ruvp=[1.0, 0.0, 0.0, 1.0], CFL=0.5)
    U = F.field_init(m, p, F.INIT_UNIFORM)
    s = O.OracleSolver(m, p)
    s.set_state(U)
    h = Lbox / n
    c = np.sqrt(1.4)
    # edge lengths are rounded through float32 in the reference
    hf = float(np.float32(h))
    df = float(np.float32(np.sqrt(2) * h))
    expect = 0.5 * (h * h / 2) / (c * np.sqrt(hf) * 4 + c * np.sqrt(df))
    assert np.isclose(s.global_dt(), expect, rtol=1e-12)


def test_isentropic_vortex_closed_form():
    m, p, U = vortex_case(16)
    xy = m.arrays()["element_xy"]
    PI = 3.1415926535897      # the reference's PI
    r2 = (xy[0] - 5) ** 2 + (xy[1] - 5) ** 2
    du = 5 / (2 * PI) * np.exp(0.5 * (1 - r2))
    T = 1 - (G - 1) * 25 / (8 * G * PI ** 2) * np.exp(1 - r2)
    rho = T ** (1 / (G - 1))
    expect = np.stack([state(r, 1 - d * (y - 5), 1 + d * (x - 5), r * t)
                       for r, d, x, y, t in zip(rho, du, xy[0], xy[1], T)], axis=1)
    assert np.allclose(U, expect, rtol=1e-13)
    assert np.array_equal(U, O.field_init(m.soa, p, F.INIT_ISENTROPIC_VORTEX, (5, 5, 5)))


def test_uniform_flow_is_preserved_by_all_boundaries():
    m = F.Mesh.generate(F.GEN_RECT_CHANNEL, 24, 12, a=3.0, b=1.0, jitter=0.1, seed=9)
    ruvp = [1.4, 2.0, 0.0, 1.0]
    for scheme in (F.SOL_ROE, F.SOL_LLF):
        for gtype in (F.GRA_GREENGAUSS, F.GRA_LEASTSQUARE):
            p = F.default_params(inf_ruvp=ruvp, inlet_ruvp=ruvp, outlet_ruvp=ruvp, flux_scheme=scheme, gradient=gtype)
            s = O.OracleSolver(m, p)
            s.set_state(F.field_init(m, p, F.INIT_UNIFORM))
            assert np.abs(s.stage_rhs()).max() < 1e-11


def test_restriction_volume_sum_quirk():
    """A bad cell takes sum(U_nb V_nb) / sum over (4 x V_nb): the volume sum is accumulated once
    per variable (Restrict.cpp:48-52), giving a quarter of the volume-weighted mean."""
    m, p, U = vortex_case(16)
    a = m.arrays()
    c = 50
    U = U.copy()
    U[0, c] = -1.0
    s = O.OracleSolver(m, p)
    s.set_state(U)
    assert s.restrict() == 1
    nb = a["element_neighbors"][:3, c]
    nb = nb[nb >= 0]
    V = a["element_volume"][nb]
    expect = (U[:, nb] * V).sum(axis=1) / (4 * V.sum())
    assert np.allclose(s.get_state()[:, c], expect, rtol=1e-14)


def test_ssp_rk3_is_third_order_in_time():
    """Same mesh, smooth flow (no limiter, LLF): the time error of SSP-RK3 falls by ~8 per
    halving of dt; a wrong stage coefficient falls to first order (~2)."""
    m = F.Mesh.generate(F.GEN_PERIODIC_SQUARE, 16, 16, b=10.0)
    T = 0.3

    def run(cfl):
        p = F.default_params(inf_ruvp=[1, 1, 1, 1], CFL=cfl, limiter=F.LIM_NONE, flux_scheme=F.SOL_LLF,
                             time_advance=F.EVO_RK3)
        s = O.OracleSolver(m, p)
        s.set_state(F.field_init(m, p, F.INIT_ISENTROPIC_VORTEX, (5, 5, 5)))
        t = 0.0
        while t < T:
            t, _ = s.step(t, T)
        return s.get_state()

    ref = run(0.05)
    e1 = np.linalg.norm(run(0.8) - ref)
    e2 = np.linalg.norm(run(0.4) - ref)
    assert e1 / e2 > 6.0, e1 / e2


def test_forward_euler_is_first_order_in_time():
    m = F.Mesh.generate(F.GEN_PERIODIC_SQUARE, 16, 16, b=10.0)
    T = 0.3

    def run(cfl):
        p = F.default_params(inf_ruvp=[1, 1, 1, 1], CFL=cfl, limiter=F.LIM_NONE, flux_scheme=F.SOL_LLF)
        s = O.OracleSolver(m, p)
        s.set_state(F.field_init(m, p, F.INIT_ISENTROPIC_VORTEX, (5, 5, 5)))
        t = 0.0
        while t < T:
            t, _ = s.step(t, T)
        return s.get_state()

    ref = run(0.02)
    r = np.linalg.norm(run(0.4) - ref) / np.linalg.norm(run(0.2) - ref)
    assert 1.6 < r < 2.6, r


def test_shocktube_first_residual_matches_reference_log(golden_dir):
    """The one number the reference's own files give for this path on its own inputs: the density
    residual (infinity norm of dU/dt) of step 1 of its GPU run of the quick_start shock tube,
    logged as 20.7896 (FVMR/working_directory/2024-10-17-142244.LOG, CDriver.cpp:115-125).  Step 1
    is the rhs of the initial state, so it pins the oracle's reconstruction, boundary ghosts and
    Roe flux on the real mesh and the case's own input.toml, to the six printed digits."""
    import json
    from cfd_solver_2d_gpu_amd import driver as D
    from helpers import toml_tree
    with open(os.path.join(os.path.dirname(__file__), "golden", "shocktube_100_reference_log.json")) as f:
        log = json.load(f)
    assert log["steps"][0] == 1
    cfg = D.config_from_tree(toml_tree("shocktube.input.toml"))
    m = O.OracleMesh.read_su2(golden_dir / "shocktube_100.su2", cfg.equation, cfg.mesh_scale_factor)
    p = cfg.params()
    s = O.OracleSolver(m, p)
    s.set_state(O.field_init(m.soa, p, cfg.initial_type, cfg.init_args()))
    s.step(0.0, cfg.max_physical_time)
    r = s.residual(F.NORM_INF)[0]
    assert abs(r / float(log["residual_rho"][0]) - 1) < 5e-6, r


def test_cylinder_oracle_reaches_the_reference_state_at_its_time(golden_dir):
    """The reference's own output of a long run: pause_cylinder_allfar_50[20000], written after
    20,000 forward-Euler steps of the quick_start cylinder (Navier-Stokes, Sutherland, M 0.5, Roe,
    Green-Gauss + Barth, CFL 2.2) at t = 0.120747.  The oracle from the case's uniform initial
    field to that time matches it to 5e-5 (rho, rhoE), 2.4e-4 (rhou) and 1e-3 (rhov, zero at the
    start) relative L2, where the flow has moved 9-40 % (rhov 100 %): a dropped term, a wrong
    sign or a wrong boundary would leave a difference of the size of the evolution.  The
    remaining difference is the reference's FP32 and its larger, racy dt (20,000 steps against
    the oracle's 21,852).  About a minute of CPU time."""
    from cfd_solver_2d_gpu_amd import driver as D
    from helpers import toml_tree
    cfg = D.config_from_tree(toml_tree("cylinder.input.toml"))
    m = O.OracleMesh.read_su2(golden_dir / "cylinder_allfar_50.su2", cfg.equation, cfg.mesh_scale_factor)
    _, Uref, tref, iref, _ = F.Mesh.read_continue(str(golden_dir / "pause_cylinder_allfar_50_20000.dat"))
    p = cfg.params()
    U0 = O.field_init(m.soa, p, cfg.initial_type, cfg.init_args())
    s = O.OracleSolver(m, p)
    s.set_state(U0)
    t, n = 0.0, 0
    while t < tref and n < 30000:
        t, _ = s.step(t, tref)
        n += 1
    assert t == tref and 20000 < n < 24000
    U = s.get_state()
    tol = [2e-4, 1e-3, 4e-3, 2e-4]
    for k in range(4):
        err = np.linalg.norm(U[k] - Uref[k]) / np.linalg.norm(Uref[k])
        moved = np.linalg.norm(U0[k] - Uref[k]) / np.linalg.norm(Uref[k])
        assert err < tol[k] and err < 0.01 * moved, (k, err, moved)


def test_viscous_flux_of_a_linear_velocity_field():
    """rho = 1, p = 1 and a linear velocity field u = a y + b x, v = c x + d y, with constant mu
    and least squares without a limiter (exact for the linear momenta): on every interior face
    whose two cells have exact gradients, the viscous part of the flux (the Euler flux minus the
    Navier-Stokes flux of the same state) is zero in the mass and (tau_xx nx + tau_xy ny) L,
    (tau_xy nx + tau_yy ny) L in the momenta, with tau_xx = mu (2 b - 2/3 (b + d)),
    tau_yy = mu (2 d - 2/3 (b + d)), tau_xy = mu (a + c) (ViscousFluxGPU.cu:139-229).  A wrong
    sign, factor or divergence term, or a transposed gradient, fails."""
    m = F.Mesh.generate(F.GEN_RECT_CHANNEL, 16, 8, a=2.0, b=1.0, jitter=0.1, seed=3, equation=F.EQ_NS)
    ar = m.arrays()
    a, b, c, d, mu = 0.3, 0.1, -0.2, 0.05, 0.01
    x, y = ar["element_xy"]
    u, v = a * y + b * x, c * x + d * y
    U = np.array([np.ones_like(x), u, v, 1.0 / 0.4 + 0.5 * (u * u + v * v)])
    flux = {}
    for eq in (F.EQ_EULER, F.EQ_NS):
        p = F.default_params(equation=eq, viscosity_model=F.MU_CONSTANT, mu0=mu, gradient=F.GRA_LEASTSQUARE,
                             limiter=F.LIM_NONE, inf_ruvp=[1.0, 0.0, 0.0, 1.0], inlet_ruvp=[1.0, 0.0, 0.0, 1.0],
                             outlet_ruvp=[1.0, 0.0, 0.0, 1.0])
        s = O.OracleSolver(m, p)
        s.set_state(U)
        s.stage_rhs()
        flux[eq] = s.edge_flux()
    Fv = flux[F.EQ_EULER] - flux[F.EQ_NS]
    L, R = ar["edge_elementL"], ar["edge_elementR"]
    nvalid = (ar["element_neighbors"][:3] >= 0).sum(axis=0)
    e = np.nonzero((R >= 0) & (nvalid[L] >= 2) & (nvalid[np.maximum(R, 0)] >= 2))[0]
    assert len(e) > 300
    nx, ny = ar["edge_normal"][0][e], ar["edge_normal"][1][e]
    ln = ar["edge_length"][e]
    txx, tyy, txy = mu * (2 * b - 2 / 3 * (b + d)), mu * (2 * d - 2 / 3 * (b + d)), mu * (a + c)
    scale = mu * ln
    assert np.abs(Fv[0][e]).max() == 0.0
    assert np.abs(Fv[1][e] - (txx * nx + txy * ny) * ln).max() < 1e-12 * scale.max()
    assert np.abs(Fv[2][e] - (txy * nx + tyy * ny) * ln).max() < 1e-12 * scale.max()


def viscous_part(m, U, **kw):
    """Viscous part of the oracle's edge fluxes (the Euler flux minus the Navier-Stokes flux of
    the same state; ViscousFluxGPU.cu:204-213 subtracts it from the convective flux) with least
    squares without a limiter, plus the interior faces whose two cells have exact gradients of
    linear fields (two or more element neighbours each)."""
    ar = m.arrays()
    ruvp = [1.0, 0.0, 0.0, 1.0]
    flux = {}
    for eq in (F.EQ_EULER, F.EQ_NS):
        p = F.default_params(equation=eq, gradient=F.GRA_LEASTSQUARE, limiter=F.LIM_NONE, inf_ruvp=ruvp,
                             inlet_ruvp=ruvp, outlet_ruvp=ruvp, **kw)
        s = O.OracleSolver(m, p)
        s.set_state(U)
        s.stage_rhs()
        flux[eq] = s.edge_flux()
    L, R = ar["edge_elementL"], ar["edge_elementR"]
    nvalid = (ar["element_neighbors"][:3] >= 0).sum(axis=0)
    e = np.nonzero((R >= 0) & (nvalid[L] >= 2) & (nvalid[np.maximum(R, 0)] >= 2))[0]
    assert len(e) > 300
    return flux[F.EQ_EULER] - flux[F.EQ_NS], e, ar


def ns_mesh():
    return F.Mesh.generate(F.GEN_RECT_CHANNEL, 16, 8, a=2.0, b=1.0, jitter=0.1, seed=3, equation=F.EQ_NS)


def test_viscous_heat_flux_of_a_fluid_at_rest():
    """rho = 1, u = v = 0 and a linear rho E: T = p / (rho R) = (gamma - 1) rho E / R is linear and
    the energy row of the viscous flux is the Fourier heat flux k grad T . n L with
    k = mu Cp / Pr, Cp = gamma R / (gamma - 1): gamma (mu / Pr) grad(rho E) . n L exactly; the
    momentum rows vanish (ViscousFluxGPU.cu:204-213).  A heat-flux coefficient of Cv instead of
    Cp is off by the factor gamma."""
    m = ns_mesh()
    x, y = m.arrays()["element_xy"]
    Ex, Ey, mu, Pr = 0.3, -0.2, 0.01, 0.72
    U = np.array([np.ones_like(x), np.zeros_like(x), np.zeros_like(x), 2.5 + Ex * x + Ey * y])
    Fv, e, ar = viscous_part(m, U, viscosity_model=F.MU_CONSTANT, mu0=mu, Pr_viscous=Pr)
    nx, ny, ln = ar["edge_normal"][0][e], ar["edge_normal"][1][e], ar["edge_length"][e]
    expect = 1.4 * mu / Pr * (Ex * nx + Ey * ny) * ln
    assert np.abs(Fv[1][e]).max() < 1e-15 and np.abs(Fv[2][e]).max() < 1e-15
    assert np.abs(Fv[3][e] - expect).max() < 1e-12 * np.abs(expect).max()


def test_viscous_energy_flux_of_a_shear_flow():
    """rho = 1, u = a y, v = c x (tau_xx = tau_yy = 0, tau_xy = mu (a + c)) and a linear rho E.
    The energy row of the viscous flux is (k grad T + tau . u) . n L with the face velocity
    u_f = (u_L + u_R) / 2 (the face state is the mean of the two cells) and, from
    T = (rho E / rho - |u|^2 / 2) / Cv, grad T = (grad(rho E) - u grad u - v grad v) / Cv =
    (Ex - c v_f, Ey - a u_f) / Cv; tau . u = (tau_xy v_f, tau_xy u_f).  Dropping either work
    term, or a sign in grad T, fails."""
    m = ns_mesh()
    x, y = m.arrays()["element_xy"]
    a, c, Ex, Ey, mu, Pr = 0.4, -0.15, 0.3, 0.25, 0.02, 0.72
    u, v = a * y, c * x
    U = np.array([np.ones_like(x), u, v, 3.0 + Ex * x + Ey * y])
    Fv, e, ar = viscous_part(m, U, viscosity_model=F.MU_CONSTANT, mu0=mu, Pr_viscous=Pr)
    L, R = ar["edge_elementL"][e], ar["edge_elementR"][e]
    uf, vf = 0.5 * (u[L] + u[R]), 0.5 * (v[L] + v[R])
    nx, ny, ln = ar["edge_normal"][0][e], ar["edge_normal"][1][e], ar["edge_length"][e]
    Rgas, g = 287.06, 1.4
    Cv = Rgas / (g - 1)
    k = mu * g * Cv / Pr
    txy = mu * (a + c)
    qx = k * (Ex - c * vf) / Cv + txy * vf
    qy = k * (Ey - a * uf) / Cv + txy * uf
    expect = (qx * nx + qy * ny) * ln
    assert np.abs(Fv[1][e] - txy * ny * ln).max() < 1e-12 * abs(txy)
    assert np.abs(Fv[2][e] - txy * nx * ln).max() < 1e-12 * abs(txy)
    assert np.abs(Fv[3][e] - expect).max() < 1e-12 * np.abs(expect).max()


@pytest.mark.parametrize("T0", [288.16, 250.0])
def test_sutherland_viscosity(T0):
    """Sutherland's law mu(T) = mu0 (T / T0)^(3/2) (T0 + S) / (T + S), S = 110.4 K
    (ViscousFluxGPU.h:20-31): mu(T0) = mu0 and the law at other temperatures, read through the
    heat flux of a fluid at rest (as above, gamma mu(T_f) / Pr grad(rho E) . n L) whose face
    temperature T_f = mean(rho E) / Cv spans about 330-530 K over the channel."""
    m = ns_mesh()
    x, y = m.arrays()["element_xy"]
    Rgas, g, mu0, Pr, S = 287.06, 1.4, 1.8e-5, 0.72, 110.4
    Cv = Rgas / (g - 1)
    Ex, Ey = 100.0 * Cv, 20.0 * Cv
    U = np.array([np.ones_like(x), np.zeros_like(x), np.zeros_like(x), 330.0 * Cv + Ex * x + Ey * y])
    Fv, e, ar = viscous_part(m, U, viscosity_model=F.MU_SUTHERLAND, mu0=mu0, T0=T0, Pr_viscous=Pr)
    L, R = ar["edge_elementL"][e], ar["edge_elementR"][e]
    Tf = 0.5 * (U[3][L] + U[3][R]) / Cv
    assert Tf.min() < 360 and Tf.max() > 500
    mu = mu0 * (Tf / T0) ** 1.5 * (T0 + S) / (Tf + S)
    nx, ny, ln = ar["edge_normal"][0][e], ar["edge_normal"][1][e], ar["edge_length"][e]
    expect = g * mu / Pr * (Ex * nx + Ey * ny) * ln
    assert np.abs(Fv[3][e] - expect).max() < 1e-12 * np.abs(expect).max()
    # at T = T0 the law gives mu0 itself
    assert mu0 * (T0 / T0) ** 1.5 * (T0 + S) / (T0 + S) == mu0


def periodic_shift_map(m, n, L):
    """For every cell of the periodic square, the cell one quad further in x and in y (same
    triangle shape, centroid + (h, h) modulo L)."""
    h = L / n
    x, y = m.arrays()["element_xy"]
    kx, ky = np.rint(3 * x / h).astype(int), np.rint(3 * y / h).astype(int)
    index = {(a, b): c for c, (a, b) in enumerate(zip(kx, ky))}
    return np.array([index[((a + 3) % (3 * n), (b + 3) % (3 * n))] for a, b in zip(kx, ky)])


def test_green_gauss_periodic_faces_follow_the_translation():
    """On the periodic square the mesh is invariant under a shift by one quad, so the Green-Gauss
    gradient of a periodic field shifted by one quad is the gradient of the field, shifted.
    Cells on the periodic boundary see their neighbour across the periodic face through the
    shift of its centroid by the face midpoint difference (GradientGPU.cu:332-339); a wrong
    shift changes the Green-Gauss weight of those faces only, which breaks the invariance."""
    n, Lbox = 12, 10.0
    m = F.Mesh.generate(F.GEN_PERIODIC_SQUARE, n, n, b=Lbox)
    x, y = m.arrays()["element_xy"]
    k = 2 * np.pi / Lbox
    U = np.array([1.0 + 0.2 * np.sin(k * x) * np.cos(k * y), 0.3 + 0.1 * np.cos(k * x),
                  -0.2 + 0.1 * np.sin(k * (x + y)), 2.5 + 0.3 * np.cos(k * x) * np.sin(k * y)])
    nxt = periodic_shift_map(m, n, Lbox)
    U2 = np.empty_like(U)
    U2[:, nxt] = U
    p = F.default_params(inf_ruvp=[1, 1, 1, 1], gradient=F.GRA_GREENGAUSS, limiter=F.LIM_NONE)
    s = O.OracleSolver(m, p)
    s.set_state(U)
    gx, gy = s.gradient()
    s.set_state(U2)
    gx2, gy2 = s.gradient()
    scale = max(np.abs(gx).max(), np.abs(gy).max())
    assert np.abs(gx2[:, nxt] - gx).max() < 1e-12 * scale
    assert np.abs(gy2[:, nxt] - gy).max() < 1e-12 * scale
    # and the weights matter: the field is not linear, so the gradients vary cell to cell
    assert np.abs(gx - gx.mean(axis=1, keepdims=True)).max() > 0.1 * scale


@pytest.mark.parametrize("K", [0.0, 3.0])
def test_venkatakrishnan_published_form(K):
    """Venkatakrishnan's limiter (AIAA 93-0880, eq. 9 with eps^2 = (K h)^3, h = sqrt(area)) per
    vertex j of a cell: with D- = grad U . (x_j - x_c) and D+ = U_max - U_c (D- > 0) or
    U_min - U_c (D- < 0), the extremes over the cell and its element neighbours,
    phi_j = [(D+^2 + eps^2) D- + 2 D-^2 D+] / [(D+^2 + 2 D-^2 + D+ D- + eps^2) D-], phi_j = 1 where
    D- = 0; the gradient is scaled by min(1, min_j phi_j).  On a random field: phi = limited /
    unlimited gradient of the oracle in every cell and variable."""
    m = F.Mesh.generate(F.GEN_RECT_CHANNEL, 14, 8, a=2.0, b=1.0, jitter=0.15, seed=21)
    ar = m.arrays()
    rng = np.random.default_rng(5)
    nc = m.num_element
    U = np.array([1.0, 0.4, -0.2, 2.5])[:, None] + 0.1 * rng.standard_normal((4, nc))
    grads = {}
    for lim in (F.LIM_NONE, F.LIM_VENKATAKRISHNAN):
        p = F.default_params(gradient=F.GRA_LEASTSQUARE, limiter=lim, venkat_K=K)
        s = O.OracleSolver(m, p)
        s.set_state(U)
        grads[lim] = s.gradient()
    gx, gy = grads[F.LIM_NONE]
    lx, ly = grads[F.LIM_VENKATAKRISHNAN]
    xc, yc = ar["element_xy"]
    nb = ar["element_neighbors"][:3]
    eps2 = (K * np.sqrt(ar["element_volume"])) ** 3
    checked = 0
    for k in range(4):
        vals = np.concatenate([U[k][None], np.where(nb >= 0, U[k][np.maximum(nb, 0)], U[k][None])])
        hi, lo = vals.max(axis=0), vals.min(axis=0)
        phi = np.ones(nc)
        for j in range(3):
            nd = ar["element_nodes"][j]
            dm = gx[k] * (ar["node_xy"][0][nd] - xc) + gy[k] * (ar["node_xy"][1][nd] - yc)
            dp = np.where(dm > 0, hi - U[k], lo - U[k])
            with np.errstate(divide="ignore", invalid="ignore"):
                pj = ((dp * dp + eps2) * dm + 2 * dm * dm * dp) / ((dp * dp + 2 * dm * dm + dp * dm + eps2) * dm)
            phi = np.minimum(phi, np.where(dm != 0, pj, 1.0))
        ok = np.abs(gx[k]) > 1e-8
        assert np.allclose(lx[k][ok] / gx[k][ok], phi[ok], rtol=1e-10, atol=1e-12)
        assert np.allclose(ly[k], gy[k] * phi, rtol=1e-10, atol=1e-14)
        checked += ok.sum()
        assert (phi < 0.9).sum() > 10    # the limiter is active in many cells
    assert checked > 3 * nc


def test_isothermal_wall_ghost():
    """No-slip isothermal wall at T_w (new; the reference declares the code, FVMR/src/global/
    Constexpr.h): the ghost state has the wall's pressure p_R = p_L and the reversed velocity, so
    the face mean velocity is zero, and T_R = 2 T_w - T_L, so the face mean temperature is T_w
    (T = p / (rho R)); the ghost temperature is held at T_w / 2 or above when T_L > 1.5 T_w."""
    Rgas, g = 287.06, 1.4
    for rho, u, v, p, Tw in [(1.2, 30.0, -12.0, 1.0e5, 300.0), (0.8, -5.0, 40.0, 9.0e4, 420.0),
                             (1.0, 1.0, 1.0, 2.0e5, 300.0)]:
        UL = state(rho, u, v, p, g)
        UR = O.isothermal(UL, Tw, Rgas, g)
        rl, rr = O.u2ruvp(UL, g), O.u2ruvp(UR, g)
        TL, TR = rl[3] / (rl[0] * Rgas), rr[3] / (rr[0] * Rgas)
        assert np.isclose(rr[3], p, rtol=1e-14)
        assert np.allclose(rr[1:3], [-u, -v], rtol=1e-13)
        if TL < 1.5 * Tw:
            assert abs(0.5 * (TL + TR) - Tw) < 1e-12 * Tw
        else:
            assert abs(TR - 0.5 * Tw) < 1e-12 * Tw
