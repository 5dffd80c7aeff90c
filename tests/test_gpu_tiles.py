"""The tile kernels (one work-group per tile of cells, gradients in LDS, one flux per face)
against the cell kernels and the oracle, through the C ABI.

Both kernel families evaluate the same source on the same values (the tile kernels recompute
centroids, midpoints and normals from the node coordinates with the host's arithmetic, bit
for bit); the compiler contracts multiply-adds of the inlined flux code differently in the
two kernels, so their fields agree to rounding: 1e-13 relative L2 here (1e-6 in FP32-mixed),
against the oracle's 1e-12.
"""
import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
import oracle as O
from conftest import rel_l2
from helpers import SoaMesh, cylinder_case, vortex_case, wedge_case

pytestmark = pytest.mark.gpu
TOL = 1e-12


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def run(m, p, U, steps, kernels, **kw):
    for k, v in kw.items():
        setattr(p, k, v)
    p.kernels = kernels
    g = F.GPUSolver2(m, p)
    g.set_state(U)
    t = 0.0
    for _ in range(steps):
        t, _ = g.iteration_device(t, 1e9)
    out = g.get_state()
    g.close()
    return out, t


CASES = {
    "vortex_gg_roe_rk3": lambda gd: vortex_case(48),
    "vortex_lsq_llf_venkat": lambda gd: vortex_case(48, gradient=F.GRA_LEASTSQUARE, scheme=F.SOL_LLF,
                                                  limiter=F.LIM_VENKATAKRISHNAN),
    "vortex_gg_nolimiter_euler": lambda gd: vortex_case(40, limiter=F.LIM_NONE, advance=F.EVO_EXPLICIT),
    "wedge_roe_rk3": lambda gd: wedge_case(96, 48, scheme=F.SOL_ROE),
    "wedge_constant_reconstruction": lambda gd: wedge_case(64, 32, scheme=F.SOL_ROE, reconstruct=F.REC_CONSTANT),
    "cylinder_ns_euler": lambda gd: cylinder_case(gd),
}


@pytest.mark.parametrize("case", list(CASES))
def test_tiles_match_cells(case, golden_dir):
    announce(f"tile vs cell kernels: {case}")
    m, p, U = CASES[case](golden_dir)
    steps = 3 if case.startswith("cylinder") else 5
    Uc, tc = run(m, p, U, steps, F.KERNELS_CELL)
    Ut, tt = run(m, p, U, steps, F.KERNELS_TILE)
    assert abs(tt / tc - 1) < 1e-13
    assert rel_l2(Ut, Uc) < 1e-13


@pytest.mark.parametrize("tile_cells", [1, 7, 61, 256])
def test_tile_sizes(tile_cells):
    announce(f"tile size {tile_cells}: periodic vortex 5 RK3 steps vs the oracle")
    m, p, U = vortex_case(32)
    Ut, tt = run(m, p, U, 5, F.KERNELS_TILE, tile_cells=tile_cells)
    o = O.OracleSolver(m, p)
    o.set_state(U)
    to = 0.0
    for _ in range(5):
        to, _ = o.step(to, 1e9)
    assert abs(tt / to - 1) < TOL
    assert rel_l2(Ut, o.get_state()) < TOL


def test_tile_orders_agree():
    announce("tile kernels under Hilbert, RCM and file order, wedge 64x32 Roe")
    m, p, U = wedge_case(64, 32, scheme=F.SOL_ROE)
    outs = [run(m, p, U, 4, F.KERNELS_TILE, renumber=r)[0]
            for r in (F.RENUMBER_HILBERT, F.RENUMBER_RCM, F.RENUMBER_NONE)]
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_tile_mixed_precision_matches_cells():
    announce("mixed precision tile vs cell kernels, wedge 64x32 Roe")
    m, p, U = wedge_case(64, 32, scheme=F.SOL_ROE, precision=F.PREC_MIXED)
    Uc, _ = run(m, p, U, 5, F.KERNELS_CELL)
    Ut, _ = run(m, p, U, 5, F.KERNELS_TILE)
    assert rel_l2(Ut, Uc) < 1e-6


def test_nonstandard_geometry_falls_back_to_cells():
    announce("caller SoA with a centroid off by one ulp: AUTO runs the cell kernels, TILE is refused")
    m, p, U = wedge_case(32, 16, scheme=F.SOL_ROE)
    arr = F.soa_arrays(m.soa)
    arr["element_xy"][0, 5] = np.nextafter(arr["element_xy"][0, 5], np.inf)
    sm = SoaMesh(arr)
    p.kernels = F.KERNELS_TILE
    with pytest.raises(F.FvmrError) as e:
        F.GPUSolver2(sm, p)
    assert "tile kernels" in str(e.value)
    Ua, _ = run(sm, p, U, 3, F.KERNELS_AUTO)
    Uc, _ = run(sm, p, U, 3, F.KERNELS_CELL)
    assert np.array_equal(Ua, Uc)


def test_autotune_returns_a_working_tile_size():
    announce("autotune: wedge 128x64 Roe, FP64 and mixed, work-group shapes and tile sizes")
    for prec in (F.PREC_FP64, F.PREC_MIXED):
        m, p, U = wedge_case(128, 64, scheme=F.SOL_ROE)
        p.precision = prec
        cand = F.autotune_candidates(prec)
        assert len({th for th, _ in cand}) == (3 if prec == F.PREC_FP64 else 2)
        th, best, ms = F.autotune(m, p, steps=5)
        assert (th, best) in cand and ms > 0
        Ud, _ = run(m, p, U, 3, F.KERNELS_TILE)
        p.tile_threads = th
        Ub, _ = run(m, p, U, 3, F.KERNELS_TILE, tile_cells=best)
        assert rel_l2(Ub, Ud) < (1e-13 if prec == F.PREC_FP64 else 1e-6)


@pytest.mark.parametrize("precision,threads", [(F.PREC_FP64, 128), (F.PREC_FP64, 512), (F.PREC_MIXED, 256)])
def test_work_group_shapes(precision, threads):
    """Every work-group shape of the tile kernels gives the default shape's results (FP64 to
    1e-13, the same arithmetic in other tiles; mixed to 1e-6) and the oracle's to 1e-12 (FP64)."""
    announce(f"tile work-group shape: {threads} threads, precision={precision}")
    m, p, U = wedge_case(96, 48, scheme=F.SOL_ROE)
    p.precision = precision
    Ud, _ = run(m, p, U, 4, F.KERNELS_TILE)
    p.tile_threads = threads
    Us, _ = run(m, p, U, 4, F.KERNELS_TILE)
    assert rel_l2(Us, Ud) < (1e-13 if precision == F.PREC_FP64 else 1e-6)
    if precision == F.PREC_FP64:
        o = O.OracleSolver(m, p)
        o.set_state(U)
        t = 0.0
        for _ in range(4):
            t, _ = o.step(t, 1e9)
        assert rel_l2(Us, o.get_state()) < 1e-12
    with pytest.raises(F.FvmrError):
        p.tile_threads = 1024
        F.GPUSolver2(m, p)
