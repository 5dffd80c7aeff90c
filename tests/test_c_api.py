"""The C ABI from C++: tests/c_api/fvmr_c_caller.cpp includes only include/fvmr.h and
tests/c_api/fvmr_cpp_mirror.cpp the C++ layer include/fvmr.hpp (GPU::GPUSolver2,
CIteration::iteration_device); both link libfvmr.so and drive it in the order the reference's
host code drives its solver.  On the CPU they must compile cleanly and the host-only part (SU2
mesh, shock-tube field) must agree with the Python path bit for bit; on the GPU both runs must
equal the Python path's."""
import os
import subprocess

import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
from cfd_solver_2d_gpu_amd.build import CPP_MIRROR, C_CALLER, build_c_caller

MASK = (1 << 64) - 1


def state_fold(U):
    """The caller's XOR fold of the state's bit patterns (variables in order, cells in order)."""
    h = 0
    for b in np.ascontiguousarray(U, dtype=np.float64).view(np.uint64).ravel().tolist():
        h ^= (b + 0x9E3779B97F4A7C15 + ((h << 6) & MASK) + (h >> 2)) & MASK
    return h


def shocktube_params():
    p = F.default_params(inlet_ruvp=[1.0, 0.0, 0.0, 1.0], outlet_ruvp=[0.125, 0.0, 0.0, 0.1],
                         inf_ruvp=[0.125, 0.0, 0.0, 0.1], CFL=0.5, flux_scheme=F.SOL_ROE,
                         reconstruct=F.REC_LINEAR, gradient=F.GRA_GREENGAUSS, limiter=F.LIM_BARTH,
                         time_advance=F.EVO_RK3)
    return p


def test_c_caller_compiles_and_matches_host_path(tmp_path, golden_dir):
    exe = build_c_caller(out=str(tmp_path / "fvmr_c_caller"))
    mesh = str(golden_dir / "shocktube_100.su2")
    r = subprocess.run([exe, mesh, "0", "--host-only"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    w = r.stdout.split()
    m = F.Mesh.read_su2(mesh)
    U = F.field_init(m, shocktube_params(), F.INIT_SHOCK_TUBE, (0.0, 0.0, 1.0, 0.0))
    assert (int(w[1]), int(w[3]), int(w[5]), int(w[7])) == (m.num_element, m.soa.num_edge, m.soa.num_node,
                                                          m.soa.num_boundary_set)
    assert int(w[9], 16) == state_fold(U)


def test_cpp_mirror_compiles(tmp_path):
    exe = build_c_caller(out=str(tmp_path / "fvmr_cpp_mirror"), name="fvmr_cpp_mirror")
    r = subprocess.run([exe, str(tmp_path / "missing.su2"), "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and r.stderr.startswith("fvmr_mesh_read_su2: ")


def test_c_caller_reports_errors(tmp_path):
    exe = build_c_caller(out=str(tmp_path / "fvmr_c_caller"))
    r = subprocess.run([exe, str(tmp_path / "missing.su2"), "0", "--host-only"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1 and r.stderr.startswith("fvmr_mesh_read_su2: ")


@pytest.mark.gpu
def test_c_caller_steps_match_python(golden_dir):
    """The prebuilt caller (built with the library) runs 20 steps on the device; the Python path
    on the same inputs gives the same t, residual and state bit for bit."""
    for exe in (C_CALLER, CPP_MIRROR):
        if not os.path.exists(exe):
            raise RuntimeError(f"{exe} is missing: build() compiles it with the library")
    mesh = str(golden_dir / "shocktube_100.su2")
    print("case: C++ callers, shocktube_100, 20 SSP-RK3 steps", flush=True)
    r = subprocess.run([C_CALLER, mesh, "20"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    w = r.stdout.split()
    rm = subprocess.run([CPP_MIRROR, mesh, "20"], capture_output=True, text=True, timeout=120)
    assert rm.returncode == 0, rm.stderr
    wm = rm.stdout.split()
    m = F.Mesh.read_su2(mesh)
    p = shocktube_params()
    with F.GPUSolver2(m, p) as g:
        g.set_state(F.field_init(m, p, F.INIT_SHOCK_TUBE, (0.0, 0.0, 1.0, 0.0)))
        t = 0.0
        for _ in range(20):
            t, _ = g.iteration_device(t, 1e9)
        res = g.updateResidualVector(F.NORM_INF)
        U = g.get_state()
        ruvp, _, _ = g.node_field()
    assert float(w[1]) == t and float(w[3]) == res[0]
    assert int(w[5], 16) == state_fold(U)
    assert wm[:6] == w[:6]
    assert float(wm[7]) == pytest.approx(float(np.mean(ruvp[0])), rel=1e-13)
