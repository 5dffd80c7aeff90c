"""Throughput of the explicit FV residual-and-update path on one MI355X.

Workload: M = 2 Euler flow over a 15 deg ramp in a [0,3]x[0,1] channel, 1024x512 quads
split into 1,048,576 triangles, interior nodes jittered by +-10 % of the local spacing
(splitmix64 seed 2024), inlet/outlet/slip walls, Roe flux, Green-Gauss gradient + Barth
limiter, linear reconstruction, SSP-RK3, CFL 0.5, FP64.  A step is one SSP-RK3 step
(3 stages); value = cells x 3 x steps / wall time of the timed steps.

Per step the tile kernels run k_tile_dt + k_dt_finish (global dt) and 3 x (k_tile +
k_restrict); k_tile is one whole RK stage (gradients, fluxes, update).  Its roofline counts
the algorithmic bytes of one RK stage, B_stage = n_c(12 s_U + 16 s_G + 3 s_X + 24) +
n_f(8 + 5 s_X) + n_v 2 s_X, against its average launch time from HIP events; the measured
HBM traffic per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated) is reported beside.

Also reported: the same run in FP32-mixed precision (speed-up and relative L2 density
error against FP64 after the same steps), the LLF flux, the cell-kernel family on the same
workload, and the CPU oracle timed on a bounded sample.
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


def run_gpu(scheme, precision, steps, warmup, profile, kernels=None):
    import cfd_solver_2d_gpu_amd as F
    m, p, U = make_case(scheme, precision, kernels)
    g = F.GPUSolver2(m, p)
    g.set_state(U)
    t = 0.0
    if warmup:
        t, _ = g.advance(warmup, t, 1e9)
    g.synchronize()
    t0 = time.perf_counter()
    t, _ = g.advance(steps, t, 1e9)
    g.synchronize()
    wall = time.perf_counter() - t0
    state = g.get_state()
    kern = {}
    if profile:
        # second pass over the same number of steps with HIP events around every launch
        g.profile(True)
        g.profile_reset()
        g.advance(steps, t, 1e9)
        g.synchronize()
        names = KERNEL_SLOTS["cell" if kernels == F.KERNELS_CELL else "tile"]
        for k, name in names.items():
            ms, n = g.profile_query(k)
            kern[name] = (ms, n)
        g.profile(False)
    g.close()
    return m, wall, state, kern


def traffic_from_profiles():
    """HBM bytes per k_tile launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    scaled by the copy-kernel calibration; committed in profiles/ (None when absent)."""
    path = os.path.join(ROOT, "profiles", "k_tile_pmc_wedge_1M_fp64.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch_fp64:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(steps):
    """The CPU oracle (single thread, FP64) on the same mesh for a bounded number of steps."""
    import cfd_solver_2d_gpu_amd as F
    import oracle as O
    m, p, U = make_case(F.SOL_ROE, F.PREC_FP64)
    o = O.OracleSolver(m, p)
    o.set_state(U)
    t = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        t, _ = o.step(t, 1e9)
    wall = time.perf_counter() - t0
    return m.num_element * 3 * steps / wall, wall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="FP64 Roe only")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (calibrates rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import cfd_solver_2d_gpu_amd as F

    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    m, wall, U64, kern = run_gpu(F.SOL_ROE, F.PREC_FP64, args.steps, args.warmup, profile=True)
    nc, nf, nv = m.num_element, m.num_edge, m.num_node
    value = nc * 3 * args.steps / wall
    ms_step = wall / args.steps * 1e3

    st_ms, st_n = kern["k_tile"]
    b_launch = stage_bytes(nc, nf, nv, 8, 8, 8)
    achieved = b_launch / (st_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles()
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_tile<double,double> (one RK stage: gradients, limiter, fluxes, update)",
                "bytes_per_launch": int(b_launch), "avg_launch_ms": round(st_ms, 5), "launches": st_n}
    path_b = path_bytes_per_step(nc, nf, nv, 8, 8, 8)
    out = {
        "metric": "cell-RK-stage updates/s + % HBM peak; FP32-mixed vs FP64 at stated L2 error",
        "value": value, "unit": "cell-RK-stage updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "wedge M=2, 1024x512x2 = 1,048,576 triangles (jitter 10%, seed 2024), "
                               "Roe, Green-Gauss + Barth, linear reconstruction, SSP-RK3, CFL 0.5, FP64",
                   "cells": nc, "edges": nf, "nodes": nv},
        "roofline": roofline,
        "path_roofline": {"bytes_per_step": path_b, "achieved_GBs": round(path_b / (ms_step * 1e-3) / 1e9, 1),
                          "frac": round(path_b / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels_ms": {k: round(v[0], 5) for k, v in kern.items()},
    }
    if not args.no_extra:
        _, wall_mx, Umx, kmx = run_gpu(F.SOL_ROE, F.PREC_MIXED, args.steps, args.warmup, profile=True)
        err = float(np.linalg.norm(Umx[0] - U64[0]) / np.linalg.norm(U64[0]))
        out["mixed"] = {"value": nc * 3 * args.steps / wall_mx, "speedup_vs_fp64": wall / wall_mx,
                        "rel_l2_density_vs_fp64": err, "steps": args.steps,
                        "kernels_ms": {k: round(v[0], 5) for k, v in kmx.items()}}
        _, wall_llf, _, _ = run_gpu(F.SOL_LLF, F.PREC_FP64, args.steps, args.warmup, profile=False)
        out["llf_fp64_value"] = nc * 3 * args.steps / wall_llf
        _, wall_cell, Ucell, kcell = run_gpu(F.SOL_ROE, F.PREC_FP64, args.steps, args.warmup, profile=True,
                                             kernels=F.KERNELS_CELL)
        out["cell_kernels"] = {"value": nc * 3 * args.steps / wall_cell,
                               "rel_l2_density_vs_tile": float(np.linalg.norm(Ucell[0] - U64[0]) / np.linalg.norm(U64[0])),
                               "kernels_ms": {k: round(v[0], 5) for k, v in kcell.items()}}
    if not args.no_cpu:
        rate, cw = cpu_baseline(args.cpu_steps)
        out["cpu_baseline"] = {"value": rate, "unit": "cell-RK-stage updates/s", "cores": 1, "kind": "port",
                               "sample": f"CPU oracle (FP64, 1 thread, g++ -O2) on the same 1,048,576-cell wedge, "
                                         f"{args.cpu_steps} SSP-RK3 steps ({cw:.1f} s)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
