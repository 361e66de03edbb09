"""GPU: the MaxH functional (MaxH_Functional, calc_adjoint.C:420-548) and
functionals placed at a step other than the final one.

The box record (largest h over the box cells after every step) is checked
bit for bit against numpy maxima over the states of the restatement
(oracle/titan_oracle.c) taken step by step; the MaxH step against numpy's
first argmax of that record.  A sweep whose weights apply to U^s* is checked
against the restatement's own forward + reverse sweep over s* steps (which
knows nothing of later steps): lambda^0 bit-identical, dJ/dphi_bed at rtol
1e-12 (partial sums ordered by tile).  Members of one ensemble may each take
their weights at a different step."""
import sys

import numpy as np
import pytest

from discrete_adjoint_amd import _lib
from discrete_adjoint_amd.solver import Ensemble, Simulation, TgError
from helpers import analytic_problem, paraboloid, small_leaf_flow_problem

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _slide(nx=96, ny=40, nsteps=120, fx=0.2):
    """A small pile sliding down a 32-degree plane: a box just below it
    fills, peaks and drains within the run."""
    L, W = nx * 5.0, ny * 5.0
    return analytic_problem(nx, ny, 5.0, slope_x_deg=32.0, slope_y_deg=4.0,
                            h=paraboloid(fx * L, 0.5 * W, 20.0, 20.0, 4.0), phi_int=33.0,
                            phi_bed=12.0, nsteps=nsteps, curv=(5e-4, -2e-4))


def _mask(pb, fx0, fx1, fy0=0.3, fy1=0.7):
    L, W = pb.nx * pb.dx, pb.ny * pb.dy
    return pb.box_weights((fx0 * L, fx1 * L, fy0 * W, fy1 * W)).reshape(3, -1)[0] != 0


def _maxh_weights(mask):
    """func_sens of the MaxH functional: (1, 0, 0) on the box cells
    (calc_adjoint.C:483-486)."""
    w = np.zeros((3, mask.size))
    w[0, mask] = 1.0
    return w


def _U(pb):
    return [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]


def _oracle_record(orc, U0, nsteps, mask):
    U = [a.copy() for a in U0]
    rec = []
    for s in range(nsteps):
        orc.step(U, s)
        rec.append(U[0][mask].max(initial=0.0))
    return np.array(rec)


def _first_max(rec):
    return (float(rec.max()), int(np.argmax(rec)) + 1) if rec.max() > 0 else (0.0, 0)


@pytest.mark.parametrize("k,graph,store_states", [(7, 0, 0), (7, 1, 0), (10, 1, 100),
                                                  (10, 0, 60)])
def test_maxh_record_step_and_sweep(oracle_lib, k, graph, store_states):
    pb = _slide()
    mask = _mask(pb, 0.25, 0.31)
    _say(f"case maxh single member 96x40 checkpoint_every={k} graph={graph} "
         f"store={store_states}")
    rec_ref = _oracle_record(oracle_lib.Oracle.from_problem(pb), _U(pb), pb.nsteps, mask)
    hmax_ref, step_ref = _first_max(rec_ref)
    assert 0 < step_ref < pb.nsteps  # the peak lies inside the run
    w = _maxh_weights(mask)
    orc = oracle_lib.Oracle.from_problem(pb)
    _, _, lam_ref, dphi_ref = orc.forward_adjoint(_U(pb), step_ref, w, want_dphi=True)
    extra = {"use_graph": graph}
    if store_states:
        extra["store_bytes"] = 24 * pb.n * store_states
    with Simulation(pb, checkpoint_every=k, **extra) as sim:
        sim.set_maxh_region(mask)
        for rep in range(2):
            sim.set_state(*_U(pb))
            sim.run()
            assert np.array_equal(sim.box_record(), rec_ref), rep
            hm, st = sim.maxh()
            assert hm[0] == hmax_ref and st[0] == step_ref, (rep, hm, st)
            lam, dphi = sim.dual_solver(w, want_dphi=True, step=int(st[0]))
            assert np.array_equal(lam, lam_ref), rep
            assert dphi == pytest.approx(dphi_ref, rel=1e-12), rep
        # the same functional kept on the device
        sim.set_functional(w, step=int(step_ref))
        lam_d, dphi_d = sim.dual_solver(None, want_dphi=True)
        assert np.array_equal(lam_d, lam_ref) and dphi_d == dphi
        # the final step is still the default
        _, _, lam_n, _ = oracle_lib.Oracle.from_problem(pb).forward_adjoint(_U(pb), pb.nsteps, w)
        lam_f, _ = sim.dual_solver(w)
        assert np.array_equal(lam_f, lam_n)


def test_maxh_single_steps_and_errors(oracle_lib):
    """tg_forward_step records too; region off, steps out of range and
    MaxH queries without a region are errors, not silent results."""
    pb = _slide(nsteps=30)
    mask = _mask(pb, 0.2, 0.3)
    _say("case maxh single steps and error paths")
    rec_ref = _oracle_record(oracle_lib.Oracle.from_problem(pb), _U(pb), 30, mask)
    with Simulation(pb, checkpoint_every=5, max_steps=64) as sim:
        with pytest.raises(TgError):
            sim.maxh()
        sim.set_maxh_region(mask)
        for _ in range(30):
            sim.step()
        assert np.array_equal(sim.box_record(), rec_ref)
        sim.set_state(*_U(pb))
        sim.run(30)
        with pytest.raises(TgError):
            sim.dual_solver(_maxh_weights(mask), step=31)
        with pytest.raises(TgError):
            sim.dual_solver(_maxh_weights(mask), step=-1)
        sim.set_maxh_region(None)
        with pytest.raises(TgError):
            sim.maxh()
        assert sim.box_record().size == 0
        # the record stays off for later runs
        sim.set_state(*_U(pb))
        sim.run(30)
        assert sim.box_record().size == 0
        # an empty box: a zero record and step 0; weights at step 0 are lambda^0
        sim.set_maxh_region(np.zeros(pb.n, bool))
        sim.set_state(*_U(pb))
        sim.run(30)
        rec = sim.box_record()
        assert rec.size == 30 and not np.any(rec)
        hm, st = sim.maxh()
        assert hm[0] == 0.0 and st[0] == 0
        w = _maxh_weights(mask)
        lam, dphi = sim.dual_solver(w, want_dphi=True, step=0)
        assert np.array_equal(lam, w) and dphi == 0.0


def _members(nsteps):
    return [_slide(96, 40, nsteps, fx=0.2), _slide(70, 34, nsteps, fx=0.3),
            _slide(130, 20, nsteps, fx=0.15), _slide(64, 64, nsteps, fx=0.4)]


@pytest.mark.parametrize("graph", [0, 1])
def test_ensemble_members_at_their_own_steps(oracle_lib, graph):
    """Each member's weights enter at its own step (the largest first, the
    others from zero when the sweep reaches them), in one sweep; the MaxH
    record per member."""
    N = 36
    mem = _members(N)
    masks = [_mask(p, 0.2, 0.4) for p in mem]
    weights = [_maxh_weights(m) for m in masks]
    _say(f"case maxh ensemble of 4, graph={graph}")
    with Ensemble(mem, precision=_lib.TG_FP64, checkpoint_every=6, use_graph=graph) as ens:
        ens.set_maxh_region(masks)
        ens.set_state(*[np.concatenate([getattr(p, f) for p in mem]) for f in ("h", "hu", "hv")])
        ens.run(N)
        hm, st = ens.maxh()
        for b, pb in enumerate(mem):
            rec_ref = _oracle_record(oracle_lib.Oracle.from_problem(pb), _U(pb), N, masks[b])
            rec = ens.box_record(b)
            assert np.array_equal(rec, rec_ref), b
            assert (hm[b], st[b]) == _first_max(rec_ref), b
        for steps in ([N, 23, 3, 0], [12, N - 1, 30, 5], [7, 7, 7, 7]):
            lam, dphi = ens.dual_solver(weights, want_dphi=True, step=steps)
            for b, pb in enumerate(mem):
                orc = oracle_lib.Oracle.from_problem(pb)
                _, _, lam_ref, dphi_ref = orc.forward_adjoint(_U(pb), steps[b], weights[b],
                                                              want_dphi=True)
                assert np.array_equal(lam[b], lam_ref), (steps, b)
                assert dphi[b] == pytest.approx(dphi_ref, rel=1e-12, abs=0.0), (steps, b)
        ens.set_functional(weights, step=[N, 23, 3, 0])
        lam_d, dphi_d = ens.dual_solver(None, want_dphi=True)
        lam_h, dphi_h = ens.dual_solver(weights, want_dphi=True, step=[N, 23, 3, 0])
        for a, b in zip(lam_d, lam_h):
            assert np.array_equal(a, b)
        assert np.array_equal(dphi_d, dphi_h)


def test_leaf_maxh(oracle_lib):
    """The adaptive-mesh path: record over a box of leaves, the sweep from
    the MaxH step and from a step inside a recompute segment."""
    pb = small_leaf_flow_problem(nsteps=30)
    _say("case maxh leaves")
    Lx, Ly = 16 * 20.0, 12 * 20.0
    L = pb.box_weights((0.4 * Lx, 0.6 * Lx, 0.3 * Ly, 0.7 * Ly)).reshape(3, -1)[0] != 0
    assert L.any()
    rec_ref = _oracle_record(oracle_lib.LeafOracle.from_problem(pb), _U(pb), pb.nsteps, L)
    hmax_ref, step_ref = _first_max(rec_ref)
    w = _maxh_weights(L)
    with Simulation(pb, checkpoint_every=4, use_graph=1) as sim:
        sim.set_maxh_region(L)
        sim.run()
        assert np.array_equal(sim.box_record(), rec_ref)
        hm, st = sim.maxh()
        assert hm[0] == hmax_ref and st[0] == step_ref
        for s in sorted({step_ref, 13}):
            orc = oracle_lib.LeafOracle.from_problem(pb)
            _, _, lam_ref, dphi_ref = orc.forward_adjoint(_U(pb), s, w, want_dphi=True)
            lam, dphi = sim.dual_solver(w, want_dphi=True, step=s)
            assert np.array_equal(lam, lam_ref), s
            assert dphi == pytest.approx(dphi_ref, rel=1e-12), s


@pytest.mark.parametrize("prec", [_lib.TG_MIXED, _lib.TG_FP32, _lib.TG_FWD64_ADJ32])
def test_step_placed_sweep_in_every_precision(prec):
    """In every precision mode, weights placed at U^s of an N-step run give
    the terminal sweep of an s-step run bit for bit (same kernels on the
    same states), member by member of an ensemble."""
    N = 30
    mem = _members(N)[:3]
    masks = [_mask(p, 0.2, 0.4) for p in mem]
    weights = [_maxh_weights(m) for m in masks]
    steps = [N, 17, 4]
    _say(f"case staged sweep precision {prec}")

    def sweep(nrun, step):
        with Ensemble(mem, precision=prec, checkpoint_every=6, use_graph=1) as ens:
            ens.set_maxh_region(masks)
            ens.run(nrun)
            rec = [ens.box_record(b) for b in range(len(mem))]
            lam, dphi = ens.dual_solver(weights, want_dphi=True, step=step)
            return rec, lam, dphi

    rec_all, lam_all, dphi_all = sweep(N, steps)
    for b, s in enumerate(steps):
        rec_s, lam_s, dphi_s = sweep(s, None)
        assert np.array_equal(rec_all[b][:s], rec_s[b]), (prec, b)
        assert np.array_equal(lam_all[b], lam_s[b]), (prec, b)
        assert dphi_all[b] == pytest.approx(dphi_s[b], rel=1e-12, abs=0.0), (prec, b)
