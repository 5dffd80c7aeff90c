"""Device arithmetic that the parity tests cover only through whole steps: the FP64 quotient and
reciprocal forms of the kernels against IEEE division, and the least-squares gradient with the
reference's inv_2x2 determinant (compat_inv2x2) against the oracle."""
import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
import oracle as O
from conftest import rel_l2
from helpers import vortex_case

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def test_device_division_equals_ieee():
    """a / b and 1 / b over 2^20 random operand pairs spanning 1e-30 .. 1e30 in both signs (far
    beyond the solver's states, volumes and lengths), exact cases and the documented special
    cases: b = 0 or +-inf give NaN (IEEE: +-inf or 0), a = 0 gives 0."""
    announce("device FP64 division vs IEEE on 2^20 operand pairs")
    rng = np.random.default_rng(2026)
    n = 1 << 20
    a = rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-30, 30, n)
    b = rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-30, 30, n)
    # solver-like operands: densities, pressures, volumes, lengths, 3, gamma - 1
    a[:8] = [1.0, 3.0, 0.4, 1.4, 101325.0, 1e-10, 2.5e-7, 7.0]
    b[:8] = [3.0, 1.0, 1.4, 0.4, 287.06, 3.0, 1e-3, 7.0]
    q, r = F.device_division(a, b)
    assert np.array_equal(q, a / b), int(np.sum(q != a / b))
    assert np.array_equal(r, 1.0 / b), int(np.sum(r != 1.0 / b))
    z, rz = F.device_division(np.array([0.0, -0.0, 2.0, 2.0, 2.0]), np.array([5.0, 5.0, 0.0, np.inf, -np.inf]))
    assert z[0] == 0.0 and z[1] == 0.0
    assert np.isnan(z[2:]).all() and np.isnan(rz[2:]).all()


@pytest.mark.parametrize("kernels", [F.KERNELS_TILE, F.KERNELS_CELL])
def test_least_squares_with_the_reference_inverse(kernels):
    """compat_inv2x2 = 1 reproduces the reference's inv_2x2, which replaces every determinant
    >= 1e-10 by -1e-10 (MatrixKernel.h:48-49): the device gradient and stage rhs of that
    least-squares path against the oracle's (the gradients are huge; the relative bound holds)."""
    announce(f"least squares with the reference inv_2x2 (compat), kernels={kernels}")
    m, p, U = vortex_case(32, gradient=F.GRA_LEASTSQUARE, limiter=F.LIM_NONE)
    p.compat_inv2x2 = 1
    p.kernels = kernels
    g = F.GPUSolver2(m, p)
    g.set_state(U)
    o = O.OracleSolver(m, p)
    o.set_state(U)
    gx, gy = g.gradient()
    ox, oy = o.gradient()
    assert np.abs(ox).max() > 1e6            # the bug's scale
    assert rel_l2(gx, ox) < 1e-12 and rel_l2(gy, oy) < 1e-12
    rg, ro = g.stage_rhs(), o.stage_rhs()     # the stage kernel's own gradient (tile or cell)
    assert rel_l2(rg, ro) < 1e-12, rel_l2(rg, ro)
    p2 = F.default_params(inf_ruvp=[1, 1, 1, 1], gradient=F.GRA_LEASTSQUARE, limiter=F.LIM_NONE)
    o2 = O.OracleSolver(m, p2)
    o2.set_state(U)
    cx, _ = o2.gradient()
    assert rel_l2(ox, cx) > 1.0               # and it differs from the correct inverse
    g.close()
