"""Steps replayed from captured HIP graphs give the fields of eager launches, replay after
replay, with eager work and host synchronisations in between."""
import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
from helpers import vortex_case, wedge_case

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


CASES = {
    "vortex_tile_fp64": lambda: vortex_case(48),
    "wedge_tile_mixed": lambda: wedge_case(64, 32, scheme=F.SOL_ROE, precision=F.PREC_MIXED),
    "wedge_cell_fp64": lambda: wedge_case(64, 32, scheme=F.SOL_ROE, kernels=F.KERNELS_CELL),
}


@pytest.mark.parametrize("case", list(CASES))
def test_graph_replays_match_eager_steps(case):
    announce(f"graph replays vs eager steps: {case}")
    m, p, U = CASES[case]()
    ge = F.GPUSolver2(m, p)
    ge.use_graphs(False)
    gg = F.GPUSolver2(m, p)
    gg.use_graphs(True)
    ge.set_state(U)
    gg.set_state(U)
    te = tg = 0.0
    for n, norm in ((7, 0), (3, F.NORM_INF), (6, 0), (4, F.NORM_2)):
        te, ie = ge.advance(n, te, 1e9, residual_norm=norm)
        tg, ig = gg.advance(n, tg, 1e9, residual_norm=norm)
        assert te == tg
        assert np.array_equal(ge.get_state(), gg.get_state())
        if norm:
            assert np.array_equal(np.array(ie.residual[:]), np.array(ig.residual[:]))
        gg.stage_rhs()              # eager work on the same stream between replays
        ge.stage_rhs()
    ge.close()
    gg.close()


def test_batch_graph_replays():
    announce("batch of 3 cases: graph replays vs eager steps")
    cs = [vortex_case(32), vortex_case(24), vortex_case(40)]
    bs = []
    for on in (False, True):
        b = F.BatchSolver([c[0] for c in cs], [c[1] for c in cs])
        b.use_graphs(on)
        for k, c in enumerate(cs):
            b.set_state(k, c[2])
        bs.append(b)
    t = [np.zeros(3), np.zeros(3)]
    for n in (6, 5, 3):
        for k in range(2):
            t[k] = bs[k].advance(n, t[k], np.full(3, 1e9))
        assert np.array_equal(t[0], t[1])
        for c in range(3):
            assert np.array_equal(bs[0].get_state(c), bs[1].get_state(c))
    for b in bs:
        b.close()
