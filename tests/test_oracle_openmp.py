"""The OpenMP build of the oracle (the benchmark's CPU baseline: -O3 -march=native -fopenmp, no
contraction) gives the plain one-thread build's results bit for bit on several thread counts:
its loops own their writes and the time-step sums run per cell in the edge order of the
sequential loop."""
import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
import oracle as O
from helpers import vortex_case, wedge_case


def cases():
    m, p, U = vortex_case(24, gradient=F.GRA_GREENGAUSS)            # periodic faces: dt pair terms
    yield "vortex GG Barth Roe", m, p, U
    m, p, U = wedge_case(32, 16, scheme=F.SOL_LLF, gradient=F.GRA_LEASTSQUARE, limiter=F.LIM_VENKATAKRISHNAN)
    yield "wedge LSQ Venkat LLF", m, p, U
    m = F.Mesh.generate(F.GEN_CYLINDER_OGRID, 32, 16, a=0.5, b=6.0, c=1.1, equation=F.EQ_NS)
    inf = F.ruvp_from_mach(0.3, 2.0)
    p = F.default_params(equation=F.EQ_NS, inf_ruvp=inf, inlet_ruvp=inf, outlet_ruvp=inf, CFL=0.5,
                         mu0=1e-3, time_advance=F.EVO_RK3)
    yield "cylinder NS Sutherland", m, p, F.field_init(m, p, F.INIT_UNIFORM)


@pytest.mark.parametrize("threads", [1, 3, 4])
def test_openmp_oracle_equals_plain_oracle(threads):
    O.fast_lib(threads)
    for name, m, p, U in cases():
        a = O.OracleSolver(m, p)
        b = O.OracleSolver(m, p, fast=True)
        a.set_state(U)
        b.set_state(U)
        assert a.global_dt() == b.global_dt(), name
        ta = tb = 0.0
        for _ in range(3):
            ta, ia = a.step(ta, 1e9)
            tb, ib = b.step(tb, 1e9)
            assert ta == tb and ia.dt == ib.dt, name
        assert np.array_equal(a.get_state(), b.get_state()), name
        assert np.array_equal(a.residual(F.NORM_2), b.residual(F.NORM_2)), name
    assert O.fast_lib().orc_set_threads(0) == threads
