"""GPU: the adaptive time loop (discrete_adjoint_amd.adaptive: a pass of
tg_host_adapt every few steps, the GPU leaf path between passes, the reverse
sweep replaying the meshes backwards) against the CPU restatement
(oracle/adapt_oracle.py).  Bar: the same sequence of meshes, final state,
dt record and lambda^0 bit for bit, dJ/dphi_bed to 1e-12; graph replay and
kernel-by-kernel launches alike."""
import sys

import numpy as np
import pytest

from helpers import box_weights_of, small_adaptive_run
from oracle import adapt_oracle

pytestmark = pytest.mark.gpu
BOX = (90.0, 170.0, 50.0, 150.0)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("graph", [1, 0])
def test_adaptive_run_vs_oracle(oracle_lib, graph):
    run = small_adaptive_run()
    run.use_graph = graph
    N = 22
    _say(f"case adaptive run {N} steps, pass every {run.adapt_every}, graph={graph}")
    w = lambda tree: box_weights_of(run, tree, BOX)
    ref = adapt_oracle.run(run, N, w)
    try:
        tree, U = run.forward(N)
        lam, dphi = run.adjoint(w(tree))
    finally:
        run.close()
    assert len(run.segments) == len(ref["passes"])
    for seg, (rtree, rops) in zip(run.segments, ref["passes"]):
        for a, b in zip(seg.tree, rtree):
            assert np.array_equal(a, b)
        if rops is not None:
            assert np.array_equal(seg.adapt.ops, rops)
    assert any(s.adapt is not None and (s.adapt.ops["kind"] == 1).any() for s in run.segments)
    assert np.array_equal(np.concatenate([s.dts for s in run.segments]), ref["dts"])
    for a, b in zip(U, ref["U"]):
        assert np.array_equal(a, b)
    assert np.array_equal(lam, ref["lam0"])
    assert dphi == pytest.approx(ref["dphi"], rel=1e-12)
