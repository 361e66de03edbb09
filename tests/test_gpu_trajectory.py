"""GPU: the states of the last forward run (tg_get_trajectory_state), the
ones SolRec::record_solution keeps (dualmesh.C:35-78), read from the stored
trajectory or recomputed from a checkpoint with the recorded dt.  Bar: every
U^s bit-identical to the restatement's (oracle/titan_oracle.c) state after
s steps; a spill file written from a GPU state equals the one written from
the restatement's state, byte for byte after decompression."""
import gzip
import sys

import numpy as np
import pytest

from discrete_adjoint_amd import _lib, mesh
from discrete_adjoint_amd.solver import Ensemble, Simulation, TgError
from helpers import small_flow_problem, small_leaf_flow_problem

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _U(pb):
    return [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]


def _oracle_states(orc, U0, nsteps):
    U = [a.copy() for a in U0]
    out = [[a.copy() for a in U]]
    for s in range(nsteps):
        orc.step(U, s)
        out.append([a.copy() for a in U])
    return out


@pytest.mark.parametrize("store_states,graph", [(0, 1), (15, 0), (60, 1)])
def test_every_state_of_the_run(oracle_lib, store_states, graph):
    pb = small_flow_problem(nsteps=40)
    _say(f"case trajectory states 40 steps store={store_states} graph={graph}")
    ref = _oracle_states(oracle_lib.Oracle.from_problem(pb), _U(pb), 40)
    extra = {"use_graph": graph}
    if store_states:
        extra["store_bytes"] = 24 * pb.n * store_states
    with Simulation(pb, checkpoint_every=6, **extra) as sim:
        sim.run(40)
        order = list(range(41))
        for rep, seq in enumerate((order, order[::-1])):
            for s in seq:
                got = sim.trajectory_state(s)
                for a, b in zip(got, ref[s]):
                    assert np.array_equal(a, b), (rep, s)
            # the reverse sweep reuses the recompute buffers: states stay right
            sim.dual_solver(np.ones((3, pb.n)))
        with pytest.raises(TgError):
            sim.trajectory_state(41)
        with pytest.raises(TgError):
            sim.trajectory_state(-1)


def test_leaf_states_and_spill_files(oracle_lib, tmp_path):
    pb = small_leaf_flow_problem(nsteps=20)
    _say("case trajectory states leaves + spill files")
    ref = _oracle_states(oracle_lib.LeafOracle.from_problem(pb), _U(pb), 20)
    keys = pb.table.key
    with Simulation(pb, checkpoint_every=4) as sim:
        sim.run(20)
        for s in (0, 3, 4, 9, 19, 20):
            got = sim.trajectory_state(s)
            for a, b in zip(got, ref[s]):
                assert np.array_equal(a, b), s
            mesh.write_solution(tmp_path / f"g{s}", keys, got)
            mesh.write_solution(tmp_path / f"o{s}", keys, ref[s])
            g = gzip.open(tmp_path / f"g{s}", "rb").read()
            assert g == gzip.open(tmp_path / f"o{s}", "rb").read(), s
            (h, hu, hv), found, unknown = mesh.read_solution(tmp_path / f"g{s}", keys)
            assert (found, unknown) == (pb.n, 0)
            wet = ref[s][0] > 0
            assert np.array_equal(h[wet], ref[s][0][wet]) and not np.any(h[~wet])
        # the whole run as the reference's per-step files, in one call
        run = tmp_path / "run"
        run.mkdir()
        sim.write_solution_files(run, keys, rank=3)
        for s in range(21):
            got = gzip.open(run / f"solution_0003_{s:08d}", "rb").read()
            mesh.write_solution(tmp_path / "o", keys, ref[s])
            assert got == gzip.open(tmp_path / "o", "rb").read(), s
        with pytest.raises(TgError):
            sim.write_solution_files(run, keys, first=0, last=22)


def test_ensemble_states(oracle_lib):
    mem = [small_flow_problem(24, 20, 18), small_flow_problem(40, 16, 18, slope=20.0)]
    _say("case trajectory states ensemble of 2")
    with Ensemble(mem, precision=_lib.TG_FP64, checkpoint_every=5) as ens:
        ens.set_state(*[np.concatenate([getattr(p, f) for p in mem]) for f in ("h", "hu", "hv")])
        ens.run(18)
        for b, pb in enumerate(mem):
            ref = _oracle_states(oracle_lib.Oracle.from_problem(pb), _U(pb), 18)
            o, n = ens.offsets[b]
            for s in (0, 7, 13, 18):
                got = ens.trajectory_state(s)
                for a, r in zip(got, ref[s]):
                    assert np.array_equal(a[o:o + n], r), (b, s)


def test_store_that_does_not_fit_falls_back_to_checkpoints():
    # a stored trajectory larger than the device (4001 states of a 2048^2
    # block, 403 GB against 288 GiB = 309 GB of HBM) cannot be allocated:
    # the run keeps checkpoints only, the failed allocation is not reported
    # as the run's error, and results equal a run that asked for no store
    from helpers import analytic_problem, paraboloid
    n, N = 2048, 4000
    pb = analytic_problem(n, n, 5.0, slope_x_deg=20.0, h=paraboloid(2000.0, 5120.0, 60.0, 50.0,
                                                                   6.0), nsteps=N)
    w = np.zeros((3, pb.n))
    w[0, :] = 1.0
    out = []
    for store in (1 << 50, 0):
        _say(f"case store fallback 2048^2 {N} steps store_bytes={store}")
        with Simulation(pb, checkpoint_every=100, max_steps=N, store_bytes=store,
                        use_graph=0) as sim:
            sim.run(N)
            U = sim.get_state()
            lam, _ = sim.dual_solver(w)
            _, stored = sim.stored_sweep_timing()
            out.append((U, lam, stored))
            # the context stays usable
            sim.set_state(pb.h, pb.hu, pb.hv)
            sim.run(10)
    (Ua, la, sa), (Ub, lb, sb) = out
    assert sa == 0 and sb == 0
    for a, b in zip(Ua, Ub):
        assert np.array_equal(a, b)
    assert np.array_equal(la, lb)
    assert Ua[0].max() > 0
