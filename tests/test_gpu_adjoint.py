"""Steady solve and discrete adjoint through the C ABI.

The transposed-Jacobian product is pinned three ways:
  * central differences of the residual (fvmr_stage_rhs) along random directions in the state,
    the freestream angle and the node coordinates, contracted with a random psi;
  * the oracle's residual (the CPU restatement of the reference path) differentiated column by
    column on a small mesh: (dr/dU)^T psi entry by entry;
  * end to end, dJ/dalpha of a converged steady state against central differences of J over
    two more converged states (h = 1e-6 rad), and against a dense solve of the adjoint system
    whose Jacobian, dJ/dU and dr/dalpha come from the oracle.
"""
import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
import oracle as O
from helpers import SoaMesh

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def naca_mesh(ni, nj, growth):
    return F.Mesh.generate(F.GEN_NACA_OGRID, ni, nj, b=20.0, c=growth, naca_digits=12)


def naca_params(limiter=F.LIM_BARTH, scheme=F.SOL_ROE, aoa=2.0, mach=0.5, cfl=1.0, **kw):
    inf = F.ruvp_from_mach(mach, aoa)
    return F.default_params(inf_ruvp=inf, inlet_ruvp=inf, outlet_ruvp=inf, CFL=cfl, flux_scheme=scheme,
                            limiter=limiter, time_advance=F.EVO_RK3, **kw)


def wall_sets(mesh):
    return [np.asarray(e, dtype=np.int32) for (name, t, e) in mesh.boundary_sets() if name == "airfoil"]


def rotate(inf, dalpha):
    """Freestream turned by dalpha (rad), same speed."""
    r, u, v, p = inf
    c, s = np.cos(dalpha), np.sin(dalpha)
    return [r, c * u - s * v, s * u + c * v, p]


def solver(mesh, p, U0=None, iters=0, tol=0.0):
    g = F.GPUSolver2(mesh, p)
    g.set_state(F.field_init(mesh, p, F.INIT_UNIFORM) if U0 is None else U0)
    if iters:
        g.steady_solve(iters, tol)
    return g


def standard_geometry_arrays(a, xy):
    """SoA arrays with the geometry recomputed in FP64 from node coordinates xy (2 x nv):
    centroid, area, midpoint, unrounded length, normal (dy, -dx)/L."""
    b = {k: v.copy() for k, v in a.items()}
    x, y = xy
    en = a["element_nodes"][:3]
    x0, x1, x2 = x[en[0]], x[en[1]], x[en[2]]
    y0, y1, y2 = y[en[0]], y[en[1]], y[en[2]]
    b["node_xy"] = np.ascontiguousarray(xy)
    b["element_xy"] = np.stack([((x0 + x1) + x2) / 3.0, ((y0 + y1) + y2) / 3.0])
    b["element_volume"] = np.abs(0.5 * (x0 * (y1 - y2) + x1 * (y2 - y0) + x2 * (y0 - y1)))
    e0, e1 = a["edge_nodes"]
    dx, dy = x[e1] - x[e0], y[e1] - y[e0]
    L = np.sqrt(dx * dx + dy * dy)
    b["edge_length"] = L
    b["edge_xy"] = np.stack([(x[e0] + x[e1]) / 2.0, (y[e0] + y[e1]) / 2.0])
    b["edge_normal"] = np.stack([dy / L, -dx / L])
    return b


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def modulated_state(m, p, amp=0.05):
    """The freestream modulated in log-polar coordinates of the centroids: every cell has a
    gradient of the size of its own scale (an O-grid's cells grow geometrically), so limiter
    branches sit away from the 0/0 of a uniform region, where Barth's and Venkatakrishnan's
    limiters have no derivative."""
    a = F.soa_arrays(m.soa)
    x, y = a["element_xy"]
    r, th = np.log(np.hypot(x, y) + 0.5), np.arctan2(y, x)
    U0 = F.field_init(m, p, F.INIT_UNIFORM)
    w = [np.sin(3 * th) * np.cos(2 * r), np.cos(2 * th + 1) * np.sin(3 * r),
         np.sin(th - 2) * np.cos(r + 1), np.cos(3 * th) * np.sin(2 * r + 1)]
    return U0 * (1 + amp * np.stack(w))


@pytest.fixture(scope="module")
def small_state():
    """NACA 0012 O-grid 48 x 16 (1,536 triangles), M 0.5, AoA 2 deg, modulated freestream."""
    m = naca_mesh(48, 16, 1.2)
    p = naca_params()
    return m, p, modulated_state(m, p)


@pytest.mark.parametrize("limiter", [F.LIM_NONE, F.LIM_BARTH, F.LIM_VENKATAKRISHNAN])
def test_vjp_matches_residual_differences(small_state, limiter):
    announce(f"adjoint VJP vs central differences of the residual, limiter={limiter}")
    m, p0, U = small_state
    p = naca_params(limiter=limiter)
    g = F.GPUSolver2(m, p)
    rng = np.random.default_rng(11)
    psi = rng.standard_normal(U.shape)
    g.set_state(U)
    out, dalpha = g.adjoint_vjp(psi, alpha=True)
    for trial in range(3):
        v = rng.standard_normal(U.shape) * np.abs(U).mean(axis=1, keepdims=True)
        eps = 1e-7
        g.set_state(U + eps * v)
        rp = g.stage_rhs()
        g.set_state(U - eps * v)
        rm = g.stage_rhs()
        fd = np.sum(psi * (rp - rm)) / (2 * eps)
        ad = np.sum(out * v)
        assert rel(fd, ad) < 1e-6, (trial, fd, ad)
    # freestream angle
    g.set_state(U)
    h = 1e-6
    inf = list(p.inf_ruvp)
    g.set_boundary_states(rotate(inf, h))
    rp = g.stage_rhs()
    g.set_boundary_states(rotate(inf, -h))
    rm = g.stage_rhs()
    g.set_boundary_states(inf)
    fd = np.sum(psi * (rp - rm)) / (2 * h)
    assert rel(fd, dalpha) < 1e-6, (fd, dalpha)
    g.close()


def test_vjp_coordinates_match_differences(small_state):
    announce("adjoint VJP in node coordinates vs central differences on FP64 geometry")
    m, p, U = small_state
    a = F.soa_arrays(m.soa)
    xy = a["node_xy"].copy()
    base = SoaMesh(standard_geometry_arrays(a, xy))
    g = F.GPUSolver2(base, p)
    g.set_state(U)
    rng = np.random.default_rng(5)
    psi = rng.standard_normal(U.shape)
    _, dxy = g.adjoint_vjp(psi, coords=True)
    g.close()
    # perturbation scaled by the local spacing (smallest adjacent edge)
    e0, e1 = a["edge_nodes"]
    L = np.hypot(xy[0][e1] - xy[0][e0], xy[1][e1] - xy[1][e0])
    hloc = np.full(xy.shape[1], np.inf)
    np.minimum.at(hloc, e0, L)
    np.minimum.at(hloc, e1, L)
    for trial in range(2):
        d = rng.standard_normal(xy.shape) * hloc
        eps = 1e-7
        r = []
        for s in (1, -1):
            mm = SoaMesh(standard_geometry_arrays(a, xy + s * eps * d))
            gg = F.GPUSolver2(mm, p)
            gg.set_state(U)
            r.append(gg.stage_rhs())
            gg.close()
        fd = np.sum(psi * (r[0] - r[1])) / (2 * eps)
        ad = np.sum(dxy * d)
        assert rel(fd, ad) < 1e-5, (trial, fd, ad)


def oracle_jacobian(m, p, U):
    """dr/dU of the oracle's residual by central differences, column by column (4 n x 4 n,
    rows and columns variable-major: index k n + c)."""
    o = O.OracleSolver(m, p)
    n = U.shape[1]
    A = np.zeros((4 * n, 4 * n))
    for c in range(n):
        for k in range(4):
            h = 1e-7 * max(abs(U[k, c]), 1e-3)
            Up = U.copy(); Up[k, c] += h
            Um = U.copy(); Um[k, c] -= h
            o.set_state(Up); rp = o.stage_rhs()
            o.set_state(Um); rm = o.stage_rhs()
            A[:, k * n + c] = ((rp - rm) / (2 * h)).ravel()
    return A


@pytest.mark.parametrize("limiter", [F.LIM_NONE, F.LIM_BARTH, F.LIM_VENKATAKRISHNAN])
def test_vjp_matches_oracle_jacobian(limiter):
    announce(f"adjoint VJP vs the oracle's Jacobian by columns (NACA 24 x 6), limiter={limiter}")
    m = naca_mesh(24, 6, 1.6)
    p = naca_params(limiter=limiter)
    U = modulated_state(m, p)
    rng = np.random.default_rng(3)
    psi = rng.standard_normal(U.shape)
    g = F.GPUSolver2(m, p)
    g.set_state(U)
    out = g.adjoint_vjp(psi)
    g.close()
    ref = (oracle_jacobian(m, p, U).T @ psi.ravel()).reshape(U.shape)
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert err < 1e-6, err


def test_dJ_dalpha_matches_dense_oracle_adjoint():
    announce("dJ/dalpha vs a dense solve of the oracle-differentiated adjoint system (NACA 32 x 8)")
    m = naca_mesh(32, 8, 1.5)
    p = naca_params(limiter=F.LIM_NONE)
    sets = wall_sets(m)
    g = solver(m, p)
    g.steady_solve(60000, 1e-14)
    res, _ = g.adjoint_solve(sets, S_ref=1.0, max_iters=60000, tol=1e-13)
    U = g.get_state()
    psi_gpu = g.adjoint_state()
    g.close()
    n = U.shape[1]
    soa = m.soa
    inf = np.array(p.inf_ruvp[:])
    edges = np.concatenate(sets)

    def J_of(U, inf_):
        _, Cp, _ = O.node_field(soa, U, p.gamma, inf)      # Cp against the handle's freestream
        cx, cy = O.surface_force(soa, Cp, [edges], 1.0)
        al = np.arctan2(inf_[2], inf_[1])
        return -np.sin(al) * cx[0] + np.cos(al) * cy[0]

    gJ = np.zeros(4 * n)
    for c in range(n):
        for k in range(4):
            h = 1e-7 * max(abs(U[k, c]), 1e-3)
            Up = U.copy(); Up[k, c] += h
            Um = U.copy(); Um[k, c] -= h
            gJ[k * n + c] = (J_of(Up, inf) - J_of(Um, inf)) / (2 * h)
    A = oracle_jacobian(m, p, U)
    psi = np.linalg.solve(A.T, -gJ)
    h = 1e-6
    r = []
    for s in (1, -1):
        q = naca_params(limiter=F.LIM_NONE)
        q.inf_ruvp[:] = rotate(inf, s * h)
        o = O.OracleSolver(m, q)
        o.set_state(U)
        r.append(o.stage_rhs().ravel())
    r_alpha = (r[0] - r[1]) / (2 * h)
    J_alpha = (J_of(U, rotate(inf, h)) - J_of(U, rotate(inf, -h))) / (2 * h)
    ref = J_alpha + psi @ r_alpha
    print(f"dense oracle dJ/da = {ref:.10f}, GPU adjoint {res.dJ_dalpha:.10f}", flush=True)
    assert rel(ref, res.dJ_dalpha) < 1e-6
    assert np.linalg.norm(psi_gpu.ravel() - psi) / np.linalg.norm(psi) < 1e-5


@pytest.mark.parametrize("limiter", [F.LIM_NONE, F.LIM_VENKATAKRISHNAN])
def test_dJ_dalpha_matches_converged_differences(limiter):
    announce(f"dJ/dalpha vs central differences of converged J, NACA 64 x 24, limiter={limiter}")
    m = naca_mesh(64, 24, 1.15)
    p = naca_params(limiter=limiter, cfl=1.0)
    sets = wall_sets(m)
    g = solver(m, p)
    it, r0, r1 = g.steady_solve(60000, 1e-14)
    print(f"primal: {it} iterations, residual {r0:.3e} -> {r1:.3e}", flush=True)
    res, _ = g.adjoint_solve(sets, S_ref=1.0, max_iters=60000, tol=1e-12)
    print(f"adjoint: {res.iters} iterations, residual {res.res0:.3e} -> {res.res:.3e}; J={res.J:.12f} "
          f"dJ/da={res.dJ_dalpha:.10f}", flush=True)
    Ustar = g.get_state()
    inf = list(p.inf_ruvp)
    h = 1e-6
    J = []
    for s in (1, -1):
        g.set_boundary_states(rotate(inf, s * h))
        g.set_state(Ustar)
        g.steady_solve(60000, 1e-14)
        cx, cy = g.surface_force(1.0, sets)
        a = np.arctan2(rotate(inf, s * h)[2], rotate(inf, s * h)[1])
        J.append(-np.sin(a) * cx.sum() + np.cos(a) * cy.sum())
    fd = (J[0] - J[1]) / (2 * h)
    print(f"FD dJ/da = {fd:.10f}, rel diff {rel(fd, res.dJ_dalpha):.2e}", flush=True)
    g.close()
    assert rel(fd, res.dJ_dalpha) < 1e-5


def test_dJ_dX_matches_converged_differences():
    announce("dJ/dX vs central differences of converged J under a smooth airfoil deformation (NACA 48 x 16)")
    m = naca_mesh(48, 16, 1.2)
    p = naca_params(limiter=F.LIM_NONE)
    a = F.soa_arrays(m.soa)
    xy = a["node_xy"].copy()
    sid = [i + 1 for i, (name, t, e) in enumerate(m.boundary_sets()) if name == "airfoil"][0]
    sets = [np.nonzero(a["edge_setID"] == sid)[0].astype(np.int32)]
    # a bump lifting the middle of the airfoil, fading out by a chord
    x, y = xy
    d = np.stack([np.zeros_like(x), 0.02 * np.exp(-(x * x + y * y) / 0.3 ** 2)])

    def J_of(xy_, U0, iters):
        g = F.GPUSolver2(SoaMesh(standard_geometry_arrays(a, xy_)), p)
        g.set_state(U0)
        g.steady_solve(iters, 1e-14)
        cx, cy = g.surface_force(1.0, sets)
        al = np.arctan2(p.inf_ruvp[2], p.inf_ruvp[1])
        out = (-np.sin(al) * cx.sum() + np.cos(al) * cy.sum(), g.get_state(), g)
        return out

    J0, Ustar, g = J_of(xy, F.field_init(m, p, F.INIT_UNIFORM), 60000)
    res, dJdX = g.adjoint_solve(sets, 1.0, 60000, 1e-13, coords=True)
    g.close()
    assert rel(res.J, J0) < 1e-12
    ad = np.sum(dJdX * d)
    h = 1e-4
    Jp = J_of(xy + h * d, Ustar, 60000)
    Jm = J_of(xy - h * d, Ustar, 60000)
    fd = (Jp[0] - Jm[0]) / (2 * h)
    Jp[2].close(); Jm[2].close()
    print(f"dJ/dX . d: adjoint {ad:.10f}, central differences {fd:.10f}, rel {rel(ad, fd):.2e}", flush=True)
    assert rel(ad, fd) < 1e-5


def test_steady_solve_tile_and_cell_kernels_agree():
    announce("steady solve: tile kernels with per-cell dt vs the cell kernels (NACA 48 x 16, 400 iterations)")
    m = naca_mesh(48, 16, 1.2)
    out = []
    for kern in (F.KERNELS_TILE, F.KERNELS_CELL):
        p = naca_params(limiter=F.LIM_VENKATAKRISHNAN, kernels=kern)
        g = solver(m, p)
        it, r0, r1 = g.steady_solve(400, 0.0)
        out.append((g.get_state(), r0, r1))
        g.close()
    (Ut, t0, t1), (Uc, c0, c1) = out
    assert rel(t0, c0) < 1e-12 and rel(t1, c1) < 1e-8
    assert np.linalg.norm(Ut - Uc) / np.linalg.norm(Uc) < 1e-12
