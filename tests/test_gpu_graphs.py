"""Steps replayed from captured HIP graphs equal eagerly launched steps bit
for bit, replay after replay, with host reads, the eager spectral filter,
the source switching off (a second graph) and a background change (graphs
dropped and captured again) between the calls."""
import numpy as np
import pytest

from sparc_amd import abi
from sparc_amd import driver as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,prec", [("tiny", abi.FP64), ("tiny", abi.FP32),
                                       ("tall", abi.FP32), ("wide", abi.FP32)])
def test_graph_replays_equal_eager_steps(name, prec):
    W = D.workload(name)
    S = D.setup_for(W)
    nsteps = 24
    eager, _ = D.make_engine(W, S, precision=prec, nsteps=nsteps, use_graphs=False)
    graph, _ = D.make_engine(W, S, precision=prec, nsteps=nsteps, use_graphs=True)
    try:
        mask, series = D.forward_source(W, nsteps + 8)
        for e in (eager, graph):
            e.set_source(abi.SRC_SERIES, mask, series, 9)  # source off after step 9
        t = 0
        for call, n in enumerate((4, 4, 5, 3, 4)):
            if call == 3:  # a background change drops the captured graphs
                for e in (eager, graph):
                    e.set_background(abi.BG_C2, S.c2[0, 0, :] if S.c2.ndim == 3 else S.c2)
            for e in (eager, graph):
                e.step(t, n)
            t += n
            assert eager.norm(abi.V_Z) == graph.norm(abi.V_Z), (call, t)
            for f in range(6):
                a, b = eager.get(abi.ARR_A, f), graph.get(abi.ARR_A, f)
                assert np.array_equal(a, b), (call, f)
            assert np.array_equal(eager.get(abi.ARR_SCR, 5), graph.get(abi.ARR_SCR, 5))
        assert np.linalg.norm(eager.get(abi.ARR_A, 5)) > 0
    finally:
        eager.close()
        graph.close()
