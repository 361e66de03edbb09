"""GPU: forward steps whose dt is not a positive finite number, against the
oracle bit for bit.  The update's short path writes a dry tile as
update_states' zero-flux formula U - dt/dx (0 - 0) - dt/dy (0 - 0) with h
clamped at 0, and skips tiles whose input and output hold +0.0 only while dt
is finite; these cases check that the formula, not a copy, is what lands:
the clock past maxtime (negative dt) and at maxtime (dt = 0), on states with
-0.0 in dry cells (a negative dt turns them into +0.0); an all-dry domain
(dt = inf: non-finite, reported); steps replayed with a negative dt."""
import sys

import numpy as np
import pytest

from discrete_adjoint_amd import _lib
from discrete_adjoint_amd.solver import Simulation, TgError
from helpers import small_flow_problem

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _negative_zeros(pb):
    """The problem with -0.0 in h, hu, hv of its dry cells."""
    dry = pb.h == 0
    for f in (pb.h, pb.hu, pb.hv):
        f[dry] = -0.0
    assert dry.sum() > pb.n // 2
    return pb


def _same(a, b):
    return all(np.array_equal(x, y) and np.array_equal(np.signbit(x), np.signbit(y))
               for x, y in zip(a, b))


@pytest.mark.parametrize("time0,maxtime", [(0.25, 0.0), (0.5, 0.5)])
def test_steps_with_dt_not_positive(oracle_lib, time0, maxtime):
    pb = _negative_zeros(small_flow_problem(70, 45, 10))
    _say(f"case dt edge: clock {time0}, maxtime {maxtime}")
    orc = oracle_lib.Oracle.from_problem(pb)
    orc.params.maxtime = maxtime
    orc.time = time0
    U = [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]
    with Simulation(pb, max_steps=4, use_graph=0) as sim:
        sim.params.maxtime = maxtime
        sim.set_clock(0, time0)
        for s in range(3):
            dt_ref, _ = orc.step(U, s)
            st = sim.step()
            assert dt_ref <= 0 and st.dt == dt_ref, s
            assert _same(sim.get_state(), U), s
    if time0 > maxtime:  # the formula turned -0.0 of dry cells into +0.0
        assert (~np.signbit(U[1][pb.h == 0])).any()


def test_all_dry_domain_is_reported(oracle_lib):
    pb = small_flow_problem(40, 30, 4)
    for f in (pb.h, pb.hu, pb.hv):
        f[:] = 0.0
    _say("case dt edge: all dry")
    orc = oracle_lib.Oracle.from_problem(pb)
    with pytest.raises(FloatingPointError):
        orc.step([pb.h.copy(), pb.hu.copy(), pb.hv.copy()], 0)
    with Simulation(pb, max_steps=4, use_graph=0) as sim:
        with pytest.raises(TgError) as e:
            sim.step()
        assert e.value.code == _lib.TG_ERR_NONFINITE


def test_replayed_negative_dt(oracle_lib):
    pb = _negative_zeros(small_flow_problem(70, 45, 10))
    _say("case dt edge: replay with a negative dt")
    dts = [0.05, -0.03, 0.04, 0.0]
    orc = oracle_lib.Oracle.from_problem(pb)
    U = [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]
    for s, dt in enumerate(dts):
        orc.step(U, s, dt)
    with Simulation(pb, max_steps=8, use_graph=0) as sim:
        sim.replay(np.array(dts))
        assert _same(sim.get_state(), U)
