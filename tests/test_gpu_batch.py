"""Batches: independent cases of uneven size on one handle (fvmr_create_batch) advance exactly
as the same cases on their own handles, with per-case boundary states, CFL, dt and time."""
import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def cases():
    """Channels and cylinder O-grids of uneven size, Mach 0.3-0.8, AoA -5..5 degrees, plus a
    periodic vortex: one shared scheme, per-case freestream and CFL."""
    rng = np.random.default_rng(11)
    out = []
    for k in range(6):
        Ma, aoa = rng.uniform(0.3, 0.8), rng.uniform(-5, 5)
        inf = F.ruvp_from_mach(Ma, aoa)
        if k % 2 == 0:
            ni = int(rng.integers(16, 64))
            m = F.Mesh.generate(F.GEN_RECT_CHANNEL, ni, ni // 2, a=2.0, b=1.0, jitter=0.1, seed=k)
        else:
            m = F.Mesh.generate(F.GEN_CYLINDER_OGRID, int(rng.integers(24, 64)), 16, a=0.5, b=8.0, c=1.1)
        p = F.default_params(inf_ruvp=inf, inlet_ruvp=inf, outlet_ruvp=inf, CFL=float(rng.uniform(0.3, 0.6)),
                             time_advance=F.EVO_RK3)
        out.append((m, p, F.field_init(m, p, F.INIT_UNIFORM)))
    m = F.Mesh.generate(F.GEN_PERIODIC_SQUARE, 40, 40, a=0.0, b=10.0)
    p = F.default_params(inf_ruvp=[1, 1, 1, 1], CFL=0.5, time_advance=F.EVO_RK3)
    out.append((m, p, F.field_init(m, p, F.INIT_ISENTROPIC_VORTEX, (5.0, 5.0, 5.0))))
    # the uniform freestream of a channel stays uniform: perturb it so that the flow moves
    for m, p, U in out[:-1]:
        U *= (1 + 0.02 * np.sin(0.37 * np.arange(U.shape[1])))[None, :]
    return out


@pytest.mark.parametrize("precision", [F.PREC_FP64, F.PREC_MIXED])
def test_batch_equals_separate_solvers(precision):
    announce(f"batch of 7 uneven cases vs separate handles, precision={precision}")
    cs = cases()
    for _, p, _ in cs:
        p.precision = precision
    b = F.BatchSolver([m for m, _, _ in cs], [p for _, p, _ in cs])
    begin = b.cases()
    assert begin[-1] == sum(m.num_element for m, _, _ in cs)
    for c, (_, _, U) in enumerate(cs):
        b.set_state(c, U)
    T = np.full(len(cs), 1e9)
    T[2] = 1e-6                                   # one case reaches its end time and stops
    t = b.advance(4, np.zeros(len(cs)), T)
    for c, (m, p, U) in enumerate(cs):
        g = F.GPUSolver2(m, p)
        g.set_state(U)
        tc, _ = g.advance(4, 0.0, T[c])
        assert tc == t[c], (c, tc, t[c])
        assert np.array_equal(g.get_state(), b.get_state(c)), c
        g.close()
    assert t[2] == 1e-6
    b.close()


def test_batch_case_against_the_oracle():
    announce("batch case 5 (cylinder O-grid) against the oracle, 3 RK3 steps")
    cs = cases()
    b = F.BatchSolver([m for m, _, _ in cs], [p for _, p, _ in cs])
    for c, (_, _, U) in enumerate(cs):
        b.set_state(c, U)
    t = b.advance(3, np.zeros(len(cs)), np.full(len(cs), 1e9))
    m, p, U = cs[5]
    o = O.OracleSolver(m, p)
    o.set_state(U)
    to = 0.0
    for _ in range(3):
        to, _ = o.step(to, 1e9)
    assert abs(t[5] / to - 1) < 1e-12
    Uo = o.get_state()
    assert np.linalg.norm(b.get_state(5) - Uo) / np.linalg.norm(Uo) < 1e-12
    b.close()


def test_batch_refuses_mixed_settings():
    announce("batch with cases that differ in a shared setting is refused")
    cs = cases()[:2]
    cs[1][1].flux_scheme = F.SOL_LLF
    with pytest.raises(F.FvmrError) as e:
        F.BatchSolver([m for m, _, _ in cs], [p for _, p, _ in cs])
    assert "shared setting" in str(e.value)


def test_batch_restriction_counts_per_case():
    """A bad cell injected into one case (not the first) is repaired in that case and counted
    there: the case's restriction count and state equal the same case on its own handle, and the
    other cases count none."""
    announce("batch: restriction of a bad cell of case 3 counted in case 3")
    cs = cases()
    k = 3
    m, p, U = cs[k]
    U = U.copy()
    U[0, U.shape[1] // 2] = -1.0                  # negative density
    cs[k] = (m, p, U)
    b = F.BatchSolver([m for m, _, _ in cs], [p for _, p, _ in cs])
    for c, (_, _, Uc) in enumerate(cs):
        b.set_state(c, Uc)
    T = np.full(len(cs), 1e-3)                     # a fixed end time: both runs clamp alike
    t = b.advance(1, np.zeros(len(cs)), T)
    counts = [b.step_info(c).n_restricted for c in range(len(cs))]
    g = F.GPUSolver2(m, p)
    g.set_state(U)
    tg, ig = g.iteration_device(0.0, T[k])
    assert ig.n_restricted >= 1
    assert counts[k] == ig.n_restricted, counts
    assert sum(counts) == counts[k], counts
    assert b.step_info(k).t == tg == t[k]
    Ub, Ug = b.get_state(k), g.get_state()
    assert np.array_equal(np.isnan(Ub), np.isnan(Ug))
    assert np.array_equal(Ub[~np.isnan(Ub)], Ug[~np.isnan(Ug)])
    # a further step: case k repairs what its own handle repairs, the other cases nothing
    t = b.advance(1, t, np.full(len(cs), 1e9))
    tg, ig = g.iteration_device(tg, 1e9)
    counts = [b.step_info(c).n_restricted for c in range(len(cs))]
    assert counts[k] == ig.n_restricted and sum(counts) == counts[k], counts
    assert t[k] == tg
    Ub, Ug = b.get_state(k), g.get_state()
    assert np.array_equal(np.isnan(Ub), np.isnan(Ug))
    assert np.array_equal(Ub[~np.isnan(Ub)], Ug[~np.isnan(Ug)])
    g.close()
    b.close()


def test_batch_restriction_deferred_within_a_call():
    """Dense cells (flagged by the first step) in two cases other than the first: four steps of
    the batch in one call, whose restriction happens in the next step's time-step kernels, equal
    four one-step calls of each case on its own handle bit for bit."""
    announce("batch: deferred restriction of dense cells in cases 2 and 5, 4 steps in one call")
    cs = cases()
    for k in (2, 5):
        m, p, U = cs[k]
        U = U.copy()
        n = U.shape[1]
        for c in (n // 3, n // 3 + 1, (2 * n) // 3):
            U[:, c] *= 400.0              # density above the allowed 300
        cs[k] = (m, p, U)
    b = F.BatchSolver([m for m, _, _ in cs], [p for _, p, _ in cs])
    for c, (_, _, Uc) in enumerate(cs):
        b.set_state(c, Uc)
    t = b.advance(4, np.zeros(len(cs)), np.full(len(cs), 1e9))
    for c, (m, p, Uc) in enumerate(cs):
        g = F.GPUSolver2(m, p)
        g.set_state(Uc)
        tg, first = g.iteration_device(0.0, 1e9)
        for _ in range(3):
            tg, ig = g.iteration_device(tg, 1e9)
        assert first.n_restricted == (3 if c in (2, 5) else 0), (c, first.n_restricted)
        assert b.step_info(c).n_restricted == ig.n_restricted
        assert t[c] == tg
        assert np.array_equal(b.get_state(c), g.get_state()), c
        g.close()
    b.close()
