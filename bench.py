"""Throughput of the explicit FV residual-and-update path on one MI355X.

Workload: M = 2 Euler flow over a 15 deg ramp in a [0,3]x[0,1] channel, 1024x512 quads
split into 1,048,576 triangles, interior nodes jittered by +-10 % of the local spacing
(splitmix64 seed 2024), inlet/outlet/slip walls, Roe flux, Green-Gauss gradient + Barth
limiter, linear reconstruction, SSP-RK3, CFL 0.5, FP64.  A step is one SSP-RK3 step
(3 stages).  After an untimed clock warm-up (1000 steps) and the --warmup steps, --windows
(at least 5) windows of exactly --steps steps are timed, each between two device
synchronisations; value = cells x 3 x steps / the median window's wall time, with the
windows' spread reported beside it.

Per step the tile kernels run k_tile_dt + k_dt_finish (global dt), 3 x k_tile and one
k_restrict; k_tile is one whole RK stage (gradients, fluxes, update).  Its roofline counts
the algorithmic bytes of one RK stage, B_stage = n_c(12 s_U + 16 s_G + 3 s_X + 24) +
n_f(8 + 5 s_X) + n_v 2 s_X, against its average launch time from HIP events on the handle's
stream in a window right after the timed ones; the measured HBM traffic per launch (rocprofv3
FETCH_SIZE/WRITE_SIZE, copy-calibrated) is reported beside.

Also reported: the same windows in FP32-mixed precision and a fixed 1000-step run of both
precisions from the same field (speed-up, relative L2/Linf of rho, rho u, rho v, rho E against
FP64), the LLF flux, the cell-kernel family, the C1 vortex (100 and 1000 steps), the reference's
quick_start Neuvillette case, SU2 ingest at 1M cells, the C3 Navier-Stokes cylinder, the C4
batch (tile size from the library's tuner, mixed-vs-FP64 errors on sampled cases), the C5
adjoint, and the CPU oracle (-O3 -march=native -fopenmp) on one thread and on all threads.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NI, NJ = 1024, 512


def make_case(scheme, precision, kernels=None):
    import cfd_solver_2d_gpu_amd as F
    m = F.Mesh.generate(F.GEN_WEDGE_CHANNEL, NI, NJ, a=3.0, b=1.0, c=0.5, d=15.0, jitter=0.1, seed=2024)
    ruvp = [1.4, 2.0, 0.0, 1.0]
    p = F.default_params(inf_ruvp=ruvp, inlet_ruvp=ruvp, outlet_ruvp=ruvp, CFL=0.5, flux_scheme=scheme,
                         gradient=F.GRA_GREENGAUSS, limiter=F.LIM_BARTH, reconstruct=F.REC_LINEAR,
                         time_advance=F.EVO_RK3, precision=precision)
    if kernels is not None:
        p.kernels = kernels
    U = F.field_init(m, p, F.INIT_UNIFORM)
    return m, p, U


def stage_bytes(nc, nf, nv, s_u, s_g, s_x):
    """Algorithmic bytes of one RK stage (gradient + limiter + fluxes + update): U_stage, U^n,
    U_new (4 s_U each), Ux/Uy written and read (16 s_G), centroid + volume (3 s_X), cell->node
    and cell->face int32 (24); per face L/R int32 (8) + normal, length, midpoint (5 s_X); per
    node xy (2 s_X)."""
    return nc * (12 * s_u + 16 * s_g + 3 * s_x + 24) + nf * (8 + 5 * s_x) + nv * 2 * s_x


def dt_bytes(nc, nf, s_u, s_x):
    """Algorithmic bytes of the global time step: state and volume per cell, per face L/R
    int32 + normal and length."""
    return nc * (4 * s_u + s_x) + nf * (8 + 3 * s_x)


def path_bytes_per_step(nc, nf, nv, s_u, s_g, s_x):
    """Algorithmic bytes of one SSP-RK3 step: 3 B_stage + B_dt."""
    return 3 * stage_bytes(nc, nf, nv, s_u, s_g, s_x) + dt_bytes(nc, nf, s_u, s_x)


KERNEL_SLOTS = {  # profile slot -> kernel name, per kernel family
    "tile": {0: "k_tile", 1: "k_tile_dt", 2: "k_dt_finish", 3: "k_restrict"},
    "cell": {0: "k_stage", 1: "k_gradient", 2: "k_dt_finish", 3: "k_restrict"},
}


CLOCK_WARMUP_STEPS = 1000


def run_gpu(scheme, precision, steps, warmup, profile, kernels=None, reps=1, clock_warmup=CLOCK_WARMUP_STEPS):
    """clock_warmup + warmup untimed steps, then `reps` timed windows of `steps` steps each (the
    device is synchronised before and after every window), then, right after them, one window
    with HIP events around every launch.  The GPU raises its clocks over tens of milliseconds of
    sustained load and lowers them within milliseconds of idling (the stage kernel takes 72 us at
    the start of a run and 58-60 us after about 30 ms), so the untimed clock warm-up comes first
    and the profiled window follows the timed ones without a pause.  Returns the mesh, the
    window wall times, the state after the timed windows and the per-kernel event times."""
    import cfd_solver_2d_gpu_amd as F
    m, p, U = make_case(scheme, precision, kernels)
    g = F.GPUSolver2(m, p)
    g.set_state(U)
    t = 0.0
    if clock_warmup:
        t, _ = g.advance(clock_warmup, t, 1e9)
    if warmup:
        t, _ = g.advance(warmup, t, 1e9)
    g.synchronize()
    walls = []
    for _ in range(reps):
        t0 = time.perf_counter()
        t, _ = g.advance(steps, t, 1e9)
        g.synchronize()
        walls.append(time.perf_counter() - t0)
    kern = {}
    if profile:
        g.profile(True)
        g.profile_reset()
        t, _ = g.advance(steps, t, 1e9)
        g.synchronize()
        names = KERNEL_SLOTS["cell" if kernels == F.KERNELS_CELL else "tile"]
        for k, name in names.items():
            ms, n = g.profile_query(k)
            kern[name] = (ms, n)
        g.profile(False)
    state = g.get_state()
    # the boundary takes and returns host arrays (caller's element order): one state each way
    t1 = time.perf_counter()
    g.set_state(state)
    g.synchronize()
    kern["_h2d_s"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    g.get_state()
    kern["_d2h_s"] = time.perf_counter() - t1
    g.close()
    return m, walls, state, kern


def window_stats(walls, steps):
    """Median and spread of the timed windows, in ms per step."""
    ms = sorted(w / steps * 1e3 for w in walls)
    med = float(np.median(ms))
    return {"count": len(ms), "steps_each": steps, "ms_per_step_sorted": [round(x, 5) for x in ms],
            "median_ms_per_step": med, "min_ms_per_step": ms[0], "max_ms_per_step": ms[-1],
            "spread_pct": round(100.0 * (ms[-1] - ms[0]) / med, 2)}


def accuracy_run(scheme, precisions, steps, kernels=None):
    """`steps` steps of the same initial field in each precision, each as one timed advance:
    (ms per step, state) per precision."""
    import cfd_solver_2d_gpu_amd as F
    out = {}
    for prec in precisions:
        m, p, U = make_case(scheme, prec, kernels)
        g = F.GPUSolver2(m, p)
        g.set_state(U)
        g.synchronize()
        t0 = time.perf_counter()
        g.advance(steps, 0.0, 1e9)
        g.synchronize()
        wall = time.perf_counter() - t0
        out[prec] = (wall / steps * 1e3, g.get_state())
        g.close()
    return out


def traffic_from_profiles(kernel="k_tile", precision="fp64"):
    """HBM bytes per launch of the stage kernel (k_tile, or the cell family's k_stage) measured
    with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), scaled by the copy-kernel calibration;
    committed in profiles/ (None when absent)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_pmc_wedge_1M_fp64.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"traffic_bytes_per_launch_{precision}:"):
            return float(line.split(":")[1])
    return None


def var_errors(U, Uref):
    """Relative L2 and Linf errors of rho, rho u, rho v, rho E of U against Uref."""
    names = ("rho", "rhou", "rhov", "rhoE")
    out = {}
    for k, n in enumerate(names):
        d = U[k] - Uref[k]
        out[n] = {"l2": float(np.linalg.norm(d) / np.linalg.norm(Uref[k])),
                  "linf": float(np.max(np.abs(d)) / np.max(np.abs(Uref[k])))}
    return out


def vortex_analytic(m, t, gamma=1.4, chi=5.0, L=10.0):
    """Isentropic vortex of FieldInitializer::setInitialIsentropicVortex (FieldInitializer.cpp:
    139-177, the reference's pi and T_inf = p_inf) advected by (1, 1) for time t on the
    periodic [0, L]^2, nearest periodic image; conserved variables at the centroids."""
    pi = 3.1415926535897
    a = m.arrays()
    x, y = a["element_xy"]
    xc = (L / 2 + t) % L
    dx = (x - xc + L / 2) % L - L / 2
    dy = (y - xc + L / 2) % L - L / 2
    r2 = dx * dx + dy * dy
    cd = np.exp(0.5 * (1.0 - r2))
    u = 1.0 + chi / (2 * pi) * cd * (-dy)
    v = 1.0 + chi / (2 * pi) * cd * dx
    T = 1.0 - (gamma - 1) * chi * chi / (8 * gamma * pi * pi) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1))
    p = rho * T
    return np.stack([rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * (u * u + v * v)])


def config_c1():
    """Periodic vortex 64x64x2 = 8192 cells, Roe, least squares + Barth, SSP-RK3, CFL 0.5, 100
    steps: FP64 field against the analytic vortex; FP32-mixed against FP64 (the FP64 field
    against the oracle is a parity test, tests/test_gpu_parity.py)."""
    import cfd_solver_2d_gpu_amd as F
    m = F.Mesh.generate(F.GEN_PERIODIC_SQUARE, 64, 64, a=0.0, b=10.0)
    res = {"cells": m.num_element}
    for steps in (100, 1000):
        outs = {}
        for prec in (F.PREC_FP64, F.PREC_MIXED):
            p = F.default_params(inf_ruvp=[1, 1, 1, 1], CFL=0.5, gradient=F.GRA_LEASTSQUARE, limiter=F.LIM_BARTH,
                                 flux_scheme=F.SOL_ROE, time_advance=F.EVO_RK3, precision=prec)
            U0 = F.field_init(m, p, F.INIT_ISENTROPIC_VORTEX, (5.0, 5.0, 5.0))
            g = F.GPUSolver2(m, p)
            g.set_state(U0)
            t, _ = g.advance(steps, 0.0, 1e9)
            outs[prec] = (t, g.get_state())
            g.close()
        t, U = outs[F.PREC_FP64]
        if steps == 100:
            res["t_end"] = t
            res["fp64_vs_analytic"] = var_errors(U, vortex_analytic(m, t))
            res["mixed_vs_fp64"] = var_errors(outs[F.PREC_MIXED][1], U)
        else:
            res["steps_1000"] = {"t_end": t, "mixed_vs_fp64": var_errors(outs[F.PREC_MIXED][1], U)}
    return res


def config_neuvillette(steps):
    """The reference's quick_start Neuvillette case (33,919 cells, Navier-Stokes with Sutherland
    viscosity, Roe, Green-Gauss + Barth, forward Euler, CFL 2.5) on its own mesh with the
    parameters its own input.toml sets, read by the driver's parser: steps per second of the
    device loop, beside the reference's published rate for the same case (213.538 steps/s of
    pure computation on an RTX 3050 Laptop GPU in FP32, its README's screenshot)."""
    import gzip
    import shutil
    import tempfile
    import tomli
    import cfd_solver_2d_gpu_amd as F
    from cfd_solver_2d_gpu_amd import driver as D
    golden = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(golden, "neuvillette.input.toml"), "rb") as f:
        cfg = D.config_from_tree(tomli.load(f))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "neuvillette.su2")
        with gzip.open(os.path.join(golden, "neuvillette.su2.gz"), "rb") as g, open(path, "wb") as o:
            shutil.copyfileobj(g, o)
        t0 = time.perf_counter()
        m = F.Mesh.read_su2(path, cfg.equation, cfg.mesh_scale_factor)
        ingest = time.perf_counter() - t0
    p = cfg.params()
    g = F.GPUSolver2(m, p)
    g.set_state(F.field_init(m, p, cfg.initial_type, cfg.init_args()))
    t, _ = g.advance(20, 0.0, cfg.max_physical_time)
    g.synchronize()
    t0 = time.perf_counter()
    t, _ = g.advance(steps, t, cfg.max_physical_time)
    g.synchronize()
    wall = time.perf_counter() - t0
    g.close()
    return {"cells": m.num_element, "steps": steps, "steps_per_s": steps / wall,
            "cell_updates_per_s": m.num_element * steps / wall, "t_end": t, "su2_ingest_s": ingest,
            "reference_steps_per_s": 213.538,
            "reference_source": "FVMR README screenshot (doc/img/image-20240625133322681.png), RTX 3050 Laptop, FP32"}


def su2_ingest():
    """Reading the 1,048,576-cell wedge back from an SU2 file: parser, quad split, hashed edge
    builder with the reference's first-appearance edge order, boundary sets and SoA export."""
    import tempfile
    import cfd_solver_2d_gpu_amd as F
    m = F.Mesh.generate(F.GEN_WEDGE_CHANNEL, NI, NJ, a=3.0, b=1.0, c=0.5, d=15.0, jitter=0.1, seed=2024)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "wedge.su2")
        m.write_su2(path)
        t0 = time.perf_counter()
        again = F.Mesh.read_su2(path)
        wall = time.perf_counter() - t0
    return {"cells": again.num_element, "edges": int(again.arrays()["edge_length"].shape[0]), "seconds": wall}


def config_c3(steps):
    """Laminar NS cylinder: O-grid 512 x 512 (524,288 cells), r from 0.5 to 30, growth 1.02,
    M = 0.2, Re = 100 (constant viscosity), isothermal wall + farfield, SSP-RK3, Green-Gauss and
    least squares; FP64 and FP32-mixed throughput and mixed-vs-FP64 errors after `steps`."""
    import cfd_solver_2d_gpu_amd as F
    m = F.Mesh.generate(F.GEN_CYLINDER_OGRID, 512, 512, a=0.5, b=30.0, c=1.02, equation=F.EQ_NS,
                        isothermal_wall=True)
    inf = F.ruvp_from_mach(0.2, 0.0)
    u_inf = float(np.hypot(inf[1], inf[2]))
    T_inf = inf[3] / (inf[0] * 287.06)
    res = {"cells": m.num_element, "steps": steps}
    for name, grad in (("green_gauss", F.GRA_GREENGAUSS), ("least_squares", F.GRA_LEASTSQUARE)):
        outs = {}
        for prec in (F.PREC_FP64, F.PREC_MIXED):
            p = F.default_params(equation=F.EQ_NS, viscosity_model=F.MU_CONSTANT, mu0=inf[0] * u_inf * 1.0 / 100.0,
                                 inf_ruvp=inf, inlet_ruvp=inf, outlet_ruvp=inf, CFL=0.5, gradient=grad,
                                 time_advance=F.EVO_RK3, precision=prec, wall_temperature=T_inf)
            g = F.GPUSolver2(m, p)
            g.set_state(F.field_init(m, p, F.INIT_UNIFORM))
            g.advance(2, 0.0, 1e9)
            g.synchronize()
            t0 = time.perf_counter()
            g.advance(steps, 0.0, 1e9)
            g.synchronize()
            wall = time.perf_counter() - t0
            outs[prec] = (m.num_element * 3 * steps / wall, g.get_state())
            g.close()
        res[name] = {"fp64_value": outs[F.PREC_FP64][0], "mixed_value": outs[F.PREC_MIXED][0],
                     "mixed_speedup": outs[F.PREC_MIXED][0] / outs[F.PREC_FP64][0],
                     "mixed_vs_fp64": var_errors(outs[F.PREC_MIXED][1], outs[F.PREC_FP64][1])}
    return res


def config_c4(ncases=256, steps=500, seed=0):
    """Batched sweep: `ncases` independent Euler cases, half jittered channels and half cylinder
    O-grids, cell counts log-uniform in [2k, 200k], Mach U[0.3, 0.8], AoA U[-5, 5] deg,
    FP32-mixed, SSP-RK3, on one batch handle; against the same cases on separate handles
    advanced one after the other (timed on the first 32 cases, 50 steps)."""
    import cfd_solver_2d_gpu_amd as F
    rng = np.random.default_rng(seed)
    meshes, params, states = [], [], []
    for k in range(ncases):
        cells = float(np.exp(rng.uniform(np.log(2e3), np.log(2e5))))
        Ma, aoa = rng.uniform(0.3, 0.8), rng.uniform(-5.0, 5.0)
        n = max(8, int(round(np.sqrt(cells))))
        if k % 2 == 0:
            m = F.Mesh.generate(F.GEN_RECT_CHANNEL, n, max(4, n // 2), a=2.0, b=1.0, jitter=0.1, seed=k)
        else:
            m = F.Mesh.generate(F.GEN_CYLINDER_OGRID, n, max(4, n // 2), a=0.5, b=10.0, c=1.05)
        inf = F.ruvp_from_mach(Ma, aoa)
        p = F.default_params(inf_ruvp=inf, inlet_ruvp=inf, outlet_ruvp=inf, CFL=0.5, time_advance=F.EVO_RK3,
                             precision=F.PREC_MIXED)
        meshes.append(m)
        params.append(p)
        states.append(F.field_init(m, p, F.INIT_UNIFORM))
    total = sum(m.num_element for m in meshes)
    # the library's tuner on the whole batch (a batch handle per candidate: a sample of the cases
    # would fit the last-level cache and favour another shape than the batch itself), then the
    # batch with the work-group shape and tile size it picks
    t0 = time.perf_counter()
    tune_cases = list(range(ncases))
    best_th, best_tc, best_ms, timings = F.autotune_batch([meshes[c] for c in tune_cases],
                                                          [params[c] for c in tune_cases], steps=10)
    tune_s = time.perf_counter() - t0
    default = F.autotune_candidates(F.PREC_MIXED)[0]   # the default work-group shape and tile size
    for q in params:
        q.tile_threads = best_th
        q.tile_cells = best_tc
    t0 = time.perf_counter()
    b = F.BatchSolver(meshes, params)
    setup = time.perf_counter() - t0
    for c, U in enumerate(states):
        b.set_state(c, U)
    t = b.advance(2, np.zeros(ncases), np.full(ncases, 1e9))
    b.synchronize()
    t0 = time.perf_counter()
    t = b.advance(steps, t, np.full(ncases, 1e9))
    b.synchronize()
    wall = time.perf_counter() - t0
    # FP32-mixed against FP64 on a sample of cases spread over the sizes: the batch's state after
    # its 2 + steps steps against FP64 on a handle of the case's own (a batch case equals its
    # separate handle bit for bit)
    sample = list(range(0, ncases, max(1, ncases // 8)))[:8]
    errs = []
    for c in sample:
        Umx = b.get_state(c)
        q = F.Params.from_buffer_copy(params[c])
        q.precision = F.PREC_FP64
        g = F.GPUSolver2(meshes[c], q)
        g.set_state(states[c])
        g.advance(2 + steps, 0.0, 1e9)
        U64 = g.get_state()
        g.close()
        e = var_errors(Umx, U64)
        errs.append({"case": c, "cells": meshes[c].num_element, **{k: v for k, v in e.items()}})
    b.profile(True)
    b.profile_reset()
    b.advance(5, t, np.full(ncases, 1e9))
    b.synchronize()
    kern = {name: round(b.profile_query(i)[0], 5) for i, name in ((0, "k_tile"), (1, "k_tile_dt"),
                                                                    (2, "k_dt_finish_cases"), (3, "k_restrict"))}
    b.close()
    sub, sub_steps = min(32, ncases), 50
    sub_cells = sum(m.num_element for m in meshes[:sub])
    handles = []
    for c in range(sub):
        g = F.GPUSolver2(meshes[c], params[c])
        g.set_state(states[c])
        g.advance(2, 0.0, 1e9)
        handles.append(g)
    handles[-1].synchronize()
    t0 = time.perf_counter()
    for g in handles:
        g.advance(sub_steps, 0.0, 1e9)
    for g in handles:
        g.synchronize()
    wall_sep = time.perf_counter() - t0
    for g in handles:
        g.close()
    return {"cases": ncases, "cells": total, "steps": steps, "dtype": "f32-mixed",
            "value": total * 3 * steps / wall, "unit": "cell-RK-stage updates/s", "ms_per_step": wall / steps * 1e3,
            "setup_s": setup, "kernels_ms": kern,
            "autotune": {"cases": len(tune_cases), "cells": sum(meshes[c].num_element for c in tune_cases),
                         "steps": 10, "tile_threads": best_th, "tile_cells": best_tc, "ms_per_step": best_ms,
                         "ms_per_step_by_threads_and_tile_cells": {f"{a}x{b}": v for (a, b), v in timings.items()},
                         "default": {"tile_threads": default[0], "tile_cells": default[1]},
                         "gain_vs_default": (timings.get(default, best_ms) / best_ms) if best_ms else None,
                         "seconds": tune_s},
            "mixed_vs_fp64_sample": {"steps": 2 + steps, "cases": errs,
                                     "max_rel_l2": {k: max(e[k]["l2"] for e in errs) for k in ("rho", "rhou", "rhov", "rhoE")}},
            "separate_handles": {"cases": sub, "cells": sub_cells, "steps": sub_steps,
                                 "value": sub_cells * 3 * sub_steps / wall_sep}}


def adjoint_product_bytes(nc, nf, nv):
    """Algorithmic bytes of one adjoint product (k_adj_faces, k_adj_vjp, k_adj_update), every array
    once, intermediates written and read: per cell the state (32), the steady state's gradients
    (64), psi in, psi of the stage start and psi out (3 x 32), cell geometry (32), face slots (24),
    vertices (12), the per-cell share written and read (2 x 32), dJ/dU (32) and the pseudo-time
    step (8); per face the L/R/BC record (16), normal and length (32), midpoint (16), the two
    sides' face adjoints written and read (2 x 64) and the two per-side shares written and read
    (2 x 64); per node its coordinates (16)."""
    return nc * (32 + 64 + 3 * 32 + 32 + 24 + 12 + 2 * 32 + 32 + 8) + nf * (16 + 32 + 16 + 2 * 64 + 2 * 64) + nv * 16


def config_c5(primal_iters=260000, adjoint_iters=150000, fd_iters=140000, h=1e-6):
    """Discrete adjoint on the NACA 0012 O-grid 512 x 384 x 2 = 393,216 triangles, farfield at 20
    chords (radial growth 1.01), M 0.5, AoA 2 deg, Euler, Roe, Green-Gauss, linear reconstruction
    without limiter (smooth subsonic flow), SSP-RK3 pseudo-time with per-cell dt at CFL 2, FP64.
    J = lift coefficient of the airfoil (pressure force normal to the freestream, chord 1).  The
    primal is converged from the freestream, the adjoint iterated from psi = 0; dJ/dalpha is
    checked against central differences (h rad) of J over two more steady solves warm-started from
    the converged state."""
    import cfd_solver_2d_gpu_amd as F
    m = F.Mesh.generate(F.GEN_NACA_OGRID, 512, 384, b=20.0, c=1.01, naca_digits=12)
    inf = list(F.ruvp_from_mach(0.5, 2.0))
    p = F.default_params(inf_ruvp=inf, inlet_ruvp=inf, outlet_ruvp=inf, CFL=2.0, flux_scheme=F.SOL_ROE,
                         limiter=F.LIM_NONE, time_advance=F.EVO_RK3)
    sets = [np.asarray(e, dtype=np.int32) for (name, t, e) in m.boundary_sets() if name == "airfoil"]
    g = F.GPUSolver2(m, p)
    g.set_state(F.field_init(m, p, F.INIT_UNIFORM))
    alpha = float(np.arctan2(inf[2], inf[1]))
    speed = float(np.hypot(inf[1], inf[2]))

    def J(a):
        cx, cy = g.surface_force(1.0, sets)
        return -np.sin(a) * cx.sum() + np.cos(a) * cy.sum()

    g.synchronize()
    t0 = time.perf_counter()
    it_p, r0, r1 = g.steady_solve(primal_iters, 1e-12)
    g.synchronize()
    t_primal = time.perf_counter() - t0
    print(f"[c5] primal: {it_p} iterations, residual {r0:.3e} -> {r1:.3e}, {t_primal:.1f} s", file=sys.stderr, flush=True)
    J0 = J(alpha)
    Ustar = g.get_state()
    t0 = time.perf_counter()
    res, _ = g.adjoint_solve(sets, 1.0, adjoint_iters, 1e-12)
    t_adj = time.perf_counter() - t0
    print(f"[c5] adjoint: {res.iters} iterations, residual {res.res0:.3e} -> {res.res:.3e}, {t_adj:.1f} s",
          file=sys.stderr, flush=True)
    Js = []
    t0 = time.perf_counter()
    for s in (1, -1):
        a = alpha + s * h
        g.set_boundary_states([inf[0], speed * np.cos(a), speed * np.sin(a), inf[3]])
        g.set_state(Ustar)
        g.steady_solve(fd_iters, 0.0)
        Js.append(J(a))
    t_fd = time.perf_counter() - t0
    g.close()
    fd = (Js[0] - Js[1]) / (2 * h)
    rel = abs(res.dJ_dalpha - fd) / abs(fd)
    print(f"[c5] dJ/dalpha adjoint {res.dJ_dalpha:.10f}, central differences {fd:.10f}, rel {rel:.2e}",
          file=sys.stderr, flush=True)
    nc = m.num_element
    bprod = adjoint_product_bytes(nc, m.num_edge, m.num_node)
    gbs = 3 * bprod / (t_adj / res.iters) / 1e9
    return {"workload": "NACA 0012 O-grid 512x384x2 = 393,216 triangles, farfield 20c, M 0.5, AoA 2 deg, Euler, "
                        "Roe, Green-Gauss, no limiter, SSP-RK3 pseudo-time CFL 2 (per-cell dt), FP64",
            "cells": nc, "J_lift": J0, "dJ_dalpha_adjoint": res.dJ_dalpha, "dJ_dalpha_central_diff": fd,
            "fd_h_rad": h, "rel_diff": rel,
            "primal": {"iterations": it_p, "residual_first": r0, "residual_last": r1, "seconds": t_primal,
                       "ms_per_iteration": t_primal / it_p * 1e3},
            "adjoint": {"iterations": res.iters, "residual_first": res.res0, "residual_last": res.res,
                        "seconds": t_adj, "ms_per_iteration": t_adj / res.iters * 1e3,
                        "cell_products_per_s": nc * 3 * res.iters / t_adj,
                        "roofline": {"bound": "hbm", "bytes_per_product": bprod, "products_per_iteration": 3,
                                     "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": round(gbs / HBM_PEAK_GBS, 4)}},
            "fd_solves": {"iterations_each": fd_iters, "seconds": t_fd}}


def cpu_baseline(threads, runs=5, steps_one=3, steps_all=10):
    """The CPU oracle (FP64, the restatement of the reference's CPU path) built for this host with
    -O3 -march=native -fopenmp (no contraction, no fast-math; its loops own their writes, so every
    thread count gives the plain build's bits) on the same mesh and initial field: after one
    warm-up step, `runs` windows of `steps_one` SSP-RK3 steps on one thread (the reference's CPU
    path is single-threaded: its OpenMP pragmas are commented out) and of `steps_all` steps on
    `threads` threads; the median window, in cell-RK-stage updates/s."""
    import cfd_solver_2d_gpu_amd as F
    import oracle as O
    m, p, U = make_case(F.SOL_ROE, F.PREC_FP64)
    t0 = time.perf_counter()
    O.fast_lib()
    build_s = time.perf_counter() - t0
    res = {}
    total = 0.0
    for name, nth, steps in (("one_thread", 1, steps_one), ("all_threads", threads, steps_all)):
        used = O.fast_lib(nth).orc_set_threads(0)
        o = O.OracleSolver(m, p, fast=True)
        o.set_state(U)
        t, _ = o.step(0.0, 1e9)
        walls = []
        for _ in range(runs):
            t0 = time.perf_counter()
            for _ in range(steps):
                t, _ = o.step(t, 1e9)
            walls.append(time.perf_counter() - t0)
        total += sum(walls)
        med = float(np.median(walls))
        res[name] = {"value": m.num_element * 3 * steps / med, "threads": used, "steps_per_window": steps,
                     "windows": runs, "median_s": med, "spread_pct": round(100 * (max(walls) - min(walls)) / med, 2)}
    res["build_s"] = build_s
    res["cpu_seconds"] = total
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--clock-warmup", type=int, default=CLOCK_WARMUP_STEPS,
                    help="untimed steps before --warmup that bring the GPU clocks up (0 for counter runs)")
    ap.add_argument("--windows", type=int, default=7,
                    help="timed windows of --steps steps each (at least 5); value is their median")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="FP64 Roe only")
    ap.add_argument("--no-c5", action="store_true", help="skip the adjoint leg (about two minutes)")
    ap.add_argument("--c5-only", action="store_true", help="run only the adjoint leg and print its JSON")
    ap.add_argument("--c5-iters", type=int, nargs=3, default=[260000, 150000, 140000],
                    metavar=("PRIMAL", "ADJOINT", "FD"), help="iteration budgets of the adjoint leg")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="arithmetic of the headline leg (mixed: FP32 flux/gradients, FP64 state; for profiling it "
                         "alone; the FP32-mixed comparison leg runs in any case)")
    ap.add_argument("--kernels", choices=("tile", "cell"), default="tile",
                    help="kernel family of the FP64 headline leg (the cell family is the fallback for meshes "
                         "whose geometry the tile kernels cannot recompute)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (calibrates rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this library has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import cfd_solver_2d_gpu_amd as F

    if args.c5_only:
        print(json.dumps({"c5_adjoint": config_c5(*args.c5_iters)}), flush=True)
        return
    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cell = args.kernels == "cell"
    mixed_head = args.precision == "mixed"
    reps = max(5, args.windows)
    m, walls, U64, kern = run_gpu(F.SOL_ROE, F.PREC_MIXED if mixed_head else F.PREC_FP64, args.steps, args.warmup,
                                  profile=True, kernels=F.KERNELS_CELL if cell else None, reps=reps,
                                  clock_warmup=args.clock_warmup)
    nc, nf, nv = m.num_element, m.num_edge, m.num_node
    win = window_stats(walls, args.steps)
    ms_step = win["median_ms_per_step"]
    value = nc * 3 / (ms_step * 1e-3)

    st_ms, st_n = kern["k_stage" if cell else "k_tile"]
    sg = 4 if mixed_head else 8
    b_launch = stage_bytes(nc, nf, nv, 8, sg, sg)
    achieved = b_launch / (st_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles("k_stage" if cell else "k_tile", args.precision)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": ("k_stage<double,double> (one RK stage after k_gradient: fluxes, update; cell family)" if cell
                           else "k_tile<%s,false,false,1> (one RK stage: gradients, limiter, fluxes, "
                                "update; Green-Gauss + Barth + Roe fixed at compile time)"
                                % ("float,float" if mixed_head else "double,double")),
                "bytes_per_launch": int(b_launch), "avg_launch_ms": round(st_ms, 5), "launches": st_n}
    path_b = path_bytes_per_step(nc, nf, nv, 8, sg, sg)
    out = {
        "metric": "cell-RK-stage updates/s + % HBM peak; FP32-mixed vs FP64 at stated L2 error",
        "value": value, "unit": "cell-RK-stage updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-mixed" if mixed_head else "f64", "data": "synthetic",
        "clock_warmup_steps": args.clock_warmup,
        "config": {"workload": "wedge M=2, 1024x512x2 = 1,048,576 triangles (jitter 10%, seed 2024), "
                               "Roe, Green-Gauss + Barth, linear reconstruction, SSP-RK3, CFL 0.5, "
                               + ("FP32-mixed" if mixed_head else "FP64"),
                   "cells": nc, "edges": nf, "nodes": nv},
        "roofline": roofline,
        # value and ms_per_step: the median of the timed windows of `steps` steps
        "windows": win,
        "path_roofline": {"bytes_per_step": path_b, "achieved_GBs": round(path_b / (ms_step * 1e-3) / 1e9, 1),
                          "frac": round(path_b / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels_ms": {k: round(v[0], 5) for k, v in kern.items() if not k.startswith("_")},
        "launches_per_step": round(sum(v[1] for k, v in kern.items() if not k.startswith("_")) / args.steps, 2),
        # host arrays across the boundary (PCIe, permutation to and from the device order):
        # the rate when the timed steps start from a host state and end in one
        "host_io": {"set_state_ms": round(kern["_h2d_s"] * 1e3, 3), "get_state_ms": round(kern["_d2h_s"] * 1e3, 3),
                    "pcie_inclusive_value": nc * 3 * args.steps / (ms_step * 1e-3 * args.steps + kern["_h2d_s"]
                                                                   + kern["_d2h_s"])},
    }
    if not args.no_extra:
        _, walls_mx, Umx, kmx = run_gpu(F.SOL_ROE, F.PREC_MIXED, args.steps, args.warmup, profile=True, reps=reps, clock_warmup=args.clock_warmup)
        win_mx = window_stats(walls_mx, args.steps)
        kmx = {k: v for k, v in kmx.items() if not k.startswith("_")}
        acc = accuracy_run(F.SOL_ROE, (F.PREC_FP64, F.PREC_MIXED), 1000)
        (ms64, A64), (msmx, Amx) = acc[F.PREC_FP64], acc[F.PREC_MIXED]
        out["mixed"] = {"value": nc * 3 / (win_mx["median_ms_per_step"] * 1e-3),
                        "speedup_vs_fp64": ms_step / win_mx["median_ms_per_step"],
                        "windows": win_mx,
                        "kernels_ms": {k: round(v[0], 5) for k, v in kmx.items()},
                        "steps_1000": {"steps": 1000, "fp64_ms_per_step": ms64, "mixed_ms_per_step": msmx,
                                       "speedup_vs_fp64": ms64 / msmx,
                                       "rel_l2_density_vs_fp64": float(np.linalg.norm(Amx[0] - A64[0])
                                                                       / np.linalg.norm(A64[0])),
                                       "errors_vs_fp64": var_errors(Amx, A64)}}
        _, walls_llf, _, _ = run_gpu(F.SOL_LLF, F.PREC_FP64, args.steps, args.warmup, profile=False, reps=reps, clock_warmup=args.clock_warmup)
        out["llf_fp64_value"] = nc * 3 / (window_stats(walls_llf, args.steps)["median_ms_per_step"] * 1e-3)
        _, walls_cell, Ucell, kcell = run_gpu(F.SOL_ROE, F.PREC_FP64, args.steps, args.warmup, profile=True,
                                              kernels=F.KERNELS_CELL, reps=reps, clock_warmup=args.clock_warmup)
        out["cell_kernels"] = {"value": nc * 3 / (window_stats(walls_cell, args.steps)["median_ms_per_step"] * 1e-3),
                               "rel_l2_density_vs_tile": float(np.linalg.norm(Ucell[0] - U64[0]) / np.linalg.norm(U64[0])),
                               "kernels_ms": {k: round(v[0], 5) for k, v in kcell.items() if not k.startswith("_")}}
    if not args.no_extra:
        out["c1_vortex"] = config_c1()
        out["quickstart_neuvillette"] = config_neuvillette(2000)
        out["su2_ingest_1M"] = su2_ingest()
        out["c3_ns_cylinder"] = config_c3(2000)
        out["c4_batch"] = config_c4()
    if not args.no_extra and not args.no_c5:
        out["c5_adjoint"] = config_c5(*args.c5_iters)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
        cb = cpu_baseline(threads)
        allc, one = cb["all_threads"], cb["one_thread"]
        out["cpu_baseline"] = {
            "value": allc["value"], "unit": "cell-RK-stage updates/s", "cores": allc["threads"], "kind": "port",
            "sample": f"CPU oracle (FP64 restatement of the reference CPU path, g++ -O3 -march=native -fopenmp, "
                      f"no contraction) on the same 1,048,576-cell wedge: median of {allc['windows']} windows of "
                      f"{allc['steps_per_window']} SSP-RK3 steps on {allc['threads']} threads "
                      f"({cb['cpu_seconds']:.1f} s of timed CPU work in all)",
            "nproc": os.cpu_count(),
            "one_thread": {"value": one["value"], "cores": 1, "windows": one["windows"],
                           "steps_per_window": one["steps_per_window"], "spread_pct": one["spread_pct"]},
            "all_threads_spread_pct": allc["spread_pct"], "build_s": round(cb["build_s"], 2)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
