"""GPU parity of the adaptive-mesh (leaf-table) path through the C ABI
against the leaf-table restatement (oracle/leaf_oracle.c).

Bar: bit-identical FP64 states, dt records and adjoints; the bed-friction
gradient and the mass diagnostic, whose summation order differs, at rtol
1e-12.  Meshes carry hanging sides of both kinds (a leaf next to a coarser
one, a leaf next to two finer ones) crossed by moving wet/dry fronts."""
import sys

import numpy as np
import pytest

from discrete_adjoint_amd import _lib, problems
from discrete_adjoint_amd.solver import Simulation, TgError
from helpers import (analytic_leaf_problem, analytic_problem, disc, small_flow_problem,
                     small_leaf_flow_problem)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _U(pb):
    return [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]


def test_leaf_forward_bit_exact(oracle_lib):
    pb = small_leaf_flow_problem(nsteps=60)
    _say(f"case leaves forward small n={pb.n} steps 60")
    orc = oracle_lib.LeafOracle.from_problem(pb)
    U = _U(pb)
    dts = orc.run(U, 60)
    with Simulation(pb) as sim:
        st = sim.run(60)
        h, hu, hv = sim.get_state()
        dt_gpu = sim.dt_record()
    assert np.array_equal(dt_gpu, dts)
    assert np.array_equal(h, U[0]) and np.array_equal(hu, U[1]) and np.array_equal(hv, U[2])
    area = pb.table.dx[:, 0] * pb.table.dx[:, 1]
    assert st.mass == pytest.approx(float(np.sum(area * U[0])), rel=1e-12)


def test_leaf_forward_dem_adaptive_bit_exact(oracle_lib):
    pb = problems.adaptive_problem(nx0=32, ny0=32, dx0=40.0, max_level=3, nsteps=150)
    _say(f"case leaves forward plane-runout DEM n={pb.n} levels {pb.meta['levels']} steps 150")
    orc = oracle_lib.LeafOracle.from_problem(pb)
    U = _U(pb)
    dts = orc.run(U, pb.nsteps)
    with Simulation(pb) as sim:
        sim.run()
        h, hu, hv = sim.get_state()
        dt_gpu = sim.dt_record()
    assert np.array_equal(dt_gpu, dts)
    assert np.array_equal(h, U[0]) and np.array_equal(hu, U[1]) and np.array_equal(hv, U[2])


def test_leaf_adjoint_step_matches_oracle(oracle_lib):
    pb = small_leaf_flow_problem(nsteps=30)
    _say("case leaves adjoint single step")
    orc = oracle_lib.LeafOracle.from_problem(pb)
    U = _U(pb)
    dts = orc.run(U, 30)
    dt = dts[-1]
    rng = np.random.default_rng(4)
    lam = rng.standard_normal((3, pb.n))
    ref, g_ref = orc.adjoint_step(U, 30, dt, lam, want_dphi=True)
    with Simulation(pb, max_steps=64) as sim:
        sim.set_state(*U)
        out, g = sim.adjoint_step_grad(30, dt, lam)
    assert np.array_equal(out, ref)
    assert g == pytest.approx(g_ref, rel=1e-12)


@pytest.mark.parametrize("k,graph,store", [(1, 0, 0), (7, 0, 0), (25, 1, 0), (10, 1, 1)])
def test_leaf_dual_solver_bit_exact(oracle_lib, k, graph, store):
    pb = small_leaf_flow_problem(nsteps=40)
    _say(f"case leaves dual_solver checkpoint_every={k} graph={graph} store={store}")
    w = pb.box_weights((130.0, 260.0, 60.0, 180.0))
    orc = oracle_lib.LeafOracle.from_problem(pb)
    _, dts, lam_ref, dphi_ref = orc.forward_adjoint(_U(pb), pb.nsteps, w, want_dphi=True)
    extra = {"use_graph": graph}
    if store:
        extra["store_bytes"] = 24 * pb.n * 15  # the last 15 states kept whole
    with Simulation(pb, checkpoint_every=k, **extra) as sim:
        sim.run()
        lam, dphi = sim.dual_solver(w, want_dphi=True)
        assert np.array_equal(sim.dt_record(), dts)
    assert np.array_equal(lam, lam_ref)
    assert dphi == pytest.approx(dphi_ref, rel=1e-12)


def test_leaf_dual_solver_dem_adaptive(oracle_lib):
    pb = problems.adaptive_problem(nx0=24, ny0=24, dx0=40.0, max_level=3, nsteps=80)
    _say(f"case leaves dual_solver plane-runout DEM n={pb.n} steps 80 checkpoint 20")
    w = pb.box_weights((250.0, 450.0, 380.0, 580.0))
    orc = oracle_lib.LeafOracle.from_problem(pb)
    UN, dts, lam_ref, dphi_ref = orc.forward_adjoint(_U(pb), pb.nsteps, w, want_dphi=True)
    assert np.sum(w.reshape(3, -1) * np.array(UN)) > 0
    with Simulation(pb, checkpoint_every=20) as sim:
        sim.run()
        lam, dphi = sim.dual_solver(w, want_dphi=True)
    assert np.array_equal(lam, lam_ref)
    assert dphi == pytest.approx(dphi_ref, rel=1e-12)


def test_equal_leaves_match_the_uniform_gpu_path():
    """A leaf table of equal leaves through the leaf kernels gives the
    uniform-block kernels' results bit for bit (two independent GPU paths)."""
    upb = small_flow_problem(nsteps=30)
    _say("case leaves of one level vs uniform path")
    L = upb.nx * upb.dx
    from helpers import paraboloid
    lpb = analytic_leaf_problem(upb.nx, upb.ny, upb.dx, 0, lambda L_, I, J: I < 0,
                                slope_x_deg=25.0, slope_y_deg=3.0,
                                h=paraboloid(0.35 * L, 0.5 * upb.ny * 10.0, 60.0, 45.0, 8.0),
                                phi_int=32.0, phi_bed=15.0, nsteps=30, curv=(1e-3, -5e-4))
    cell = lpb.table.order
    w = upb.box_weights((120.0, 200.0, 40.0, 160.0))
    with Simulation(upb, checkpoint_every=10) as sim:
        sim.run()
        Uu = sim.get_state()
        lu, _ = sim.dual_solver(w)
    with Simulation(lpb, checkpoint_every=10) as sim:
        sim.run()
        Ul = sim.get_state()
        ll, _ = sim.dual_solver(w.reshape(3, -1)[:, cell])
    for k in range(3):
        assert np.array_equal(Uu[k][cell], Ul[k])
    assert np.array_equal(lu[:, cell], ll)


def test_leaf_lake_at_rest_on_gpu():
    pb = analytic_leaf_problem(20, 16, 20.0, 3, disc(10.0, 8.0, 3.0),
                               h=lambda X, Y: np.full(X.shape, 2.0), nsteps=10)
    _say(f"case leaves lake at rest n={pb.n}")
    with Simulation(pb) as sim:
        sim.run()
        h, hu, hv = sim.get_state()
    assert np.array_equal(h, pb.h) and not hu.any() and not hv.any()


def test_leaf_tables_are_checked():
    pb = small_leaf_flow_problem(nsteps=5)
    _say("case leaves table checks")
    t = pb.table
    # a neighbour that does not list the leaf back
    bad = pb.table.nbr.copy()
    e = int(np.flatnonzero((t.nkind[:, 1] == 0) & (t.nkind[:, 5] == -2))[0])
    f = int(t.nbr[e, 1])
    g = int(t.nbr[f, 1]) if t.nbr[f, 1] >= 0 else int(t.nbr[f, 0])
    bad[e, 1] = bad[e, 5] = g
    import copy
    pb2 = copy.copy(pb)
    pb2.table = copy.copy(t)
    pb2.table.nbr = bad
    with pytest.raises(TgError) as err:
        Simulation(pb2)
    assert err.value.code == _lib.TG_ERR_ARG
    # the leaf path is built for FP64 only
    with pytest.raises(TgError) as err:
        Simulation(pb, precision=_lib.TG_MIXED)
    assert err.value.code == _lib.TG_ERR_UNSUPPORTED


def test_leaf_short_paths_are_taken_and_exact(oracle_lib):
    """Dry leaves (update), void leaves (reverse step) and the dry kact pass
    run their short paths on this mesh -- the full-path counters stay below
    the leaf count -- and the instrumented run is bit-identical to the
    restatement, which has no short paths."""
    pb = small_leaf_flow_problem(nsteps=40)
    _say(f"case leaves short paths n={pb.n} steps 40")
    w = pb.box_weights((130.0, 260.0, 60.0, 180.0))
    orc = oracle_lib.LeafOracle.from_problem(pb)
    _, _, lam_ref, dphi_ref = orc.forward_adjoint(_U(pb), pb.nsteps, w, want_dphi=True)
    with Simulation(pb, checkpoint_every=10) as sim:
        sim.instrument(True)
        sim.run()
        lam, dphi = sim.dual_solver(w, want_dphi=True)
        ms, launches, full, short = sim.instrumentation()
    assert np.array_equal(lam, lam_ref)
    assert dphi == pytest.approx(dphi_ref, rel=1e-12)
    for k in range(3):
        assert launches[k] > 0
        assert 0 < full[k] < launches[k] * pb.n


@pytest.mark.parametrize("graph,store", [(0, 0), (1, 1)])
def test_leaf_marks_repeated_sweeps_bit_exact(oracle_lib, graph, store):
    """Chunks of +0.0 leaves are skipped by every leaf kernel once the zero
    marks know them; a second sweep on the same context starts from the
    marks the first left in every buffer.  Both stay bit-identical to the
    restatement."""
    pb = problems.adaptive_problem(nx0=24, ny0=24, dx0=40.0, max_level=3, nsteps=60)
    _say(f"case leaves marks n={pb.n} graph={graph} store={store}")
    w = pb.box_weights((250.0, 450.0, 380.0, 580.0))
    orc = oracle_lib.LeafOracle.from_problem(pb)
    UN, dts, lam_ref, dphi_ref = orc.forward_adjoint(_U(pb), pb.nsteps, w, want_dphi=True)
    extra = {"use_graph": graph}
    if store:
        extra["store_bytes"] = 32 * pb.n * 30
    with Simulation(pb, checkpoint_every=9, **extra) as sim:
        for rep in range(2):
            sim.set_state(*_U(pb))
            sim.run()
            h, hu, hv = sim.get_state()
            lam, dphi = sim.dual_solver(w, want_dphi=True)
            assert np.array_equal(sim.dt_record(), dts), rep
            assert np.array_equal(h, UN[0]) and np.array_equal(hu, UN[1]), rep
            assert np.array_equal(hv, UN[2]), rep
            assert np.array_equal(lam, lam_ref), rep
            assert dphi == pytest.approx(dphi_ref, rel=1e-12), rep
