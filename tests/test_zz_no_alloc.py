"""fvmr_surface_force_noalloc (the allocation-free variant of fvmr_surface_force, staging in a
fixed device array of the library) against fvmr_surface_force through the C ABI: the same
coefficients bit for bit on two calls in a row, the second with fewer sets and edges, on two
handles of different meshes; sizes beyond the array are refused with a message."""
import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
from helpers import cylinder_case

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def test_noalloc_surface_force_equals_surface_force(golden_dir):
    announce("surface force without allocation vs fvmr_surface_force, cylinder continue-file state")
    m, p, U = cylinder_case(golden_dir)
    g = F.GPUSolver2(m, p)
    g.set_state(U)
    sets = [s[2] for s in m.boundary_sets()]
    assert len(sets) == 2 and all(len(e) > 10 for e in sets)
    # first call: every set, every edge
    cx, cy = g.surface_force(1.0, sets)
    nx, ny = g.surface_force_noalloc(1.0, sets)
    assert np.array_equal(cx, nx) and np.array_equal(cy, ny), (cx, nx, cy, ny)
    assert np.abs(cx).max() > 0 and np.abs(cy).max() > 0
    # second call, smaller: one set, the first part of its edges reversed, another S_ref
    small = [np.ascontiguousarray(sets[1][: len(sets[1]) // 3][::-1])]
    cx2, cy2 = g.surface_force(0.7, small)
    nx2, ny2 = g.surface_force_noalloc(0.7, small)
    assert np.array_equal(cx2, nx2) and np.array_equal(cy2, ny2), (cx2, nx2, cy2, ny2)
    # after steps the forces follow the device state
    g.advance(3, 0.0, 1e9)
    cx3, cy3 = g.surface_force(1.0, sets)
    nx3, ny3 = g.surface_force_noalloc(1.0, sets)
    assert np.array_equal(cx3, nx3) and np.array_equal(cy3, ny3)
    assert not np.array_equal(cx3, cx)
    g.close()

    announce("surface force without allocation on a second handle (generated O-grid, perturbed state)")
    m2 = F.Mesh.generate(F.GEN_CYLINDER_OGRID, 48, 24, a=0.5, b=8.0, c=1.08)
    p2 = F.default_params(inf_ruvp=F.ruvp_from_mach(0.4, 3.0))
    U2 = F.field_init(m2, p2, F.INIT_UNIFORM)
    rng = np.random.default_rng(11)
    U2 = U2 * (1.0 + 0.05 * rng.standard_normal(U2.shape))
    g2 = F.GPUSolver2(m2, p2)
    g2.set_state(U2)
    sets2 = [s[2] for s in m2.boundary_sets()]
    for subset in (sets2, [sets2[0][:5]]):
        a = g2.surface_force(2.0, subset)
        b = g2.surface_force_noalloc(2.0, subset)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    g2.close()


def test_noalloc_surface_force_refuses_what_the_array_cannot_hold():
    announce("surface force without allocation: too many sets or edges, and interior edges, are refused")
    m = F.Mesh.generate(F.GEN_CYLINDER_OGRID, 32, 16, a=0.5, b=6.0, c=1.1)
    p = F.default_params(inf_ruvp=[1.0, 0.5, 0.0, 1.0])
    g = F.GPUSolver2(m, p)
    g.set_state(F.field_init(m, p, F.INIT_UNIFORM))
    wall = [s[2] for s in m.boundary_sets()][0]
    with pytest.raises(F.FvmrError) as e:
        g.surface_force_noalloc(1.0, [wall[:1]] * 257)
    assert "at most 256 sets" in str(e.value)
    with pytest.raises(F.FvmrError) as e:
        g.surface_force_noalloc(1.0, [np.resize(wall, 131073)])
    assert "at most 131072 edges" in str(e.value)
    interior = int(np.nonzero(m.arrays()["edge_setID"] == -1)[0][0])
    with pytest.raises(F.FvmrError) as e:
        g.surface_force_noalloc(1.0, [[interior]])
    assert "not in a boundary set" in str(e.value)
    # the largest call the array takes still works and equals the allocating entry point
    full = [np.resize(wall, 131072 // 256)] * 256
    a = g.surface_force(1.0, full)
    b = g.surface_force_noalloc(1.0, full)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    g.close()
