"""GPU: the zero marks of the FP64 uniform path (tiles known to hold +0.0
are skipped by the coefficient, update and adjoint passes) change the work,
not the result.  Bar: states, dt records and adjoints bit-identical to the
restatement (oracle/titan_oracle.c, which has no marks) and across tile
shapes, sweeps and re-runs; the bed-friction gradient at rtol 1e-12 (its
partial sums are ordered by tile).  Members are sized so that mark blocks
(64 x 16 cells) are cut by the member edge, and the flows leave most
blocks dry."""
import sys

import numpy as np
import pytest

from discrete_adjoint_amd import _lib
from discrete_adjoint_amd.solver import Ensemble, Simulation
from helpers import analytic_problem, paraboloid

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pile_problem(nx, ny, fx, fy, nsteps, slope=20.0, phi_bed=16.0, a=45.0, b=35.0, H=6.0):
    """A small paraboloid pile at (fx, fy) of the block on an inclined
    plane: the rest of the block stays +0.0."""
    L, W = nx * 5.0, ny * 5.0
    return analytic_problem(nx, ny, 5.0, slope_x_deg=slope, slope_y_deg=4.0,
                            h=paraboloid(fx * L, fy * W, a, b, H), phi_int=33.0,
                            phi_bed=phi_bed, nsteps=nsteps, curv=(5e-4, -2e-4))


def _box(pb, fx0, fx1, fy0, fy1):
    L, W = pb.nx * pb.dx, pb.ny * pb.dy
    return pb.box_weights((fx0 * L, fx1 * L, fy0 * W, fy1 * W))


def _U(pb):
    return [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]


@pytest.mark.parametrize("k,graph,store", [(7, 0, 0), (7, 1, 0), (10, 1, 1)])
def test_marked_sweep_matches_oracle(oracle_lib, k, graph, store):
    pb = _pile_problem(200, 150, 0.2, 0.3, 60)
    _say(f"case marks single member 200x150 checkpoint_every={k} graph={graph} store={store}")
    w = _box(pb, 0.25, 0.6, 0.1, 0.6)
    orc = oracle_lib.Oracle.from_problem(pb)
    UN, dts, lam_ref, dphi_ref = orc.forward_adjoint(_U(pb), pb.nsteps, w, want_dphi=True)
    assert np.sum(w.reshape(3, -1) * np.array(UN)) > 0
    extra = {"use_graph": graph}
    if store:
        extra["store_bytes"] = 24 * pb.n * 20
    with Simulation(pb, checkpoint_every=k, **extra) as sim:
        for rep in range(2):  # the second sweep starts from marks the first one left
            sim.set_state(*_U(pb))
            sim.run()
            h, hu, hv = sim.get_state()
            lam, dphi = sim.dual_solver(w, want_dphi=True)
            assert np.array_equal(sim.dt_record(), dts), rep
            assert np.array_equal(h, UN[0]) and np.array_equal(hu, UN[1]), rep
            assert np.array_equal(hv, UN[2]), rep
            assert np.array_equal(lam, lam_ref), rep
            assert dphi == pytest.approx(dphi_ref, rel=1e-12), rep


def test_marks_skip_work_and_keep_results(oracle_lib):
    """The instrumented sweep: some cells of every pass are skipped (neither
    full nor short path), and the results stay bit-identical."""
    pb = _pile_problem(256, 128, 0.15, 0.5, 40)
    _say("case marks skipped-cell counters 256x128")
    w = _box(pb, 0.1, 0.5, 0.2, 0.8)
    orc = oracle_lib.Oracle.from_problem(pb)
    _, dts, lam_ref, dphi_ref = orc.forward_adjoint(_U(pb), pb.nsteps, w, want_dphi=True)
    with Simulation(pb, checkpoint_every=8) as sim:
        sim.instrument(True)
        sim.run()
        lam, dphi = sim.dual_solver(w, want_dphi=True)
        ms, launches, full, short = sim.instrumentation()
        sim.instrument(False)
    assert np.array_equal(lam, lam_ref)
    assert dphi == pytest.approx(dphi_ref, rel=1e-12)
    for i in range(3):
        cells = launches[i] * pb.n
        assert full[i] > 0 and short[i] >= 0
        assert full[i] + short[i] < cells // 2, (i, full[i], short[i], cells)


def _members(nsteps):
    return [_pile_problem(130, 70, 0.3, 0.4, nsteps),
            _pile_problem(64, 16, 0.5, 0.5, nsteps, a=20.0, b=15.0, H=3.0),
            _pile_problem(300, 40, 0.1, 0.5, nsteps, slope=24.0, phi_bed=14.0),
            _pile_problem(96, 200, 0.5, 0.15, nsteps, slope=8.0)]


def _ens_run(mem, weights, sel=None, reps=1):
    out = []
    with Ensemble(mem, precision=_lib.TG_FP64, checkpoint_every=6, use_graph=1) as ens:
        if sel is not None:
            ens.select_tiles(sel)
        for _ in range(reps):
            ens.set_state(*[np.concatenate([getattr(p, f) for p in mem])
                            for f in ("h", "hu", "hv")])
            ens.run(mem[0].nsteps)
            st = [ens.member_state(b) for b in range(len(mem))]
            dts = [ens.member_dt_record(b) for b in range(len(mem))]
            lam, dphi = ens.dual_solver(weights, want_dphi=True)
            out.append((st, dts, lam, np.asarray(dphi, dtype=float)))
    return out


@pytest.mark.parametrize("sel", [(1, 3, 2), (3, 0, 1), (2, 2, 3)])
def test_marks_with_mixed_tile_shapes(oracle_lib, sel):
    """Different tile shapes per pass write and read the same marks; members
    whose sides are not multiples of the mark block."""
    mem = _members(30)
    _say(f"case marks ensemble of 4, tile variants {sel}")
    weights = [_box(p, 0.2, 0.8, 0.2, 0.8).reshape(3, -1) for p in mem]
    runs = _ens_run(mem, weights, sel=sel, reps=2)
    for b, pb in enumerate(mem):
        orc = oracle_lib.Oracle.from_problem(pb)
        UN, dts, lam_ref, dphi_ref = orc.forward_adjoint(_U(pb), pb.nsteps, weights[b],
                                                         want_dphi=True)
        for st, dt, lam, dphi in runs:
            assert np.array_equal(dt[b], dts), (sel, b)
            for x, y in zip(st[b], UN):
                assert np.array_equal(x, y), (sel, b)
            assert np.array_equal(lam[b], lam_ref), (sel, b)
            assert dphi[b] == pytest.approx(dphi_ref, rel=1e-12), (sel, b)


def test_marks_across_selection_changes(oracle_lib):
    """Sweeps on one context with the tile selection changed between them,
    and the autotuner's timing launches (which write the scratch buffers in
    other shapes) in between: the single-tile skips read a mark byte only in
    the shape that wrote it."""
    mem = _members(24)
    _say("case marks across tile selection changes")
    weights = [_box(p, 0.2, 0.8, 0.2, 0.8).reshape(3, -1) for p in mem]
    refs = []
    for b, pb in enumerate(mem):
        orc = oracle_lib.Oracle.from_problem(pb)
        refs.append(orc.forward_adjoint(_U(pb), pb.nsteps, weights[b], want_dphi=True))
    h0 = [np.concatenate([getattr(p, f) for p in mem]) for f in ("h", "hu", "hv")]
    with Ensemble(mem, precision=_lib.TG_FP64, checkpoint_every=6, use_graph=1) as ens:
        for sel in [(3, 3, 3), (1, 2, 3), None, (0, 3, 2), (2, 2, 2), (3, 3, 3)]:
            if sel is None:
                ens.autotune(reps=2)
            else:
                ens.select_tiles(sel)
            ens.set_state(*h0)
            ens.run(mem[0].nsteps)
            lam, dphi = ens.dual_solver(weights, want_dphi=True)
            for b, (UN, dts, lam_ref, dphi_ref) in enumerate(refs):
                assert np.array_equal(ens.member_dt_record(b), dts), (sel, b)
                for x, y in zip(ens.member_state(b), UN):
                    assert np.array_equal(x, y), (sel, b)
                assert np.array_equal(lam[b], lam_ref), (sel, b)
                assert dphi[b] == pytest.approx(dphi_ref, rel=1e-12), (sel, b)


def test_single_steps_after_a_marked_run(oracle_lib):
    """tg_forward_step / tg_adjoint_step_grad after a marked sweep: the
    single-step paths see the same marks discipline (buffers swapped, uploads
    clear their marks)."""
    pb = _pile_problem(160, 48, 0.2, 0.5, 20)
    _say("case marks single steps after a sweep")
    orc = oracle_lib.Oracle.from_problem(pb)
    U = _U(pb)
    dts = orc.run(U, 20)
    rng = np.random.default_rng(3)
    lam = rng.standard_normal((3, pb.n)) * (U[0] > 0)
    ref, g_ref = orc.adjoint_step(U, 20, dts[-1], lam, want_dphi=True)
    with Simulation(pb, checkpoint_every=5, max_steps=64) as sim:
        sim.run(15)
        sim.dual_solver(_box(pb, 0.1, 0.4, 0.2, 0.8))
        sim.set_state(*_U(pb))
        for _ in range(20):
            sim.step()
        h, hu, hv = sim.get_state()
        assert np.array_equal(sim.dt_record(), dts)
        assert np.array_equal(h, U[0]) and np.array_equal(hu, U[1]) and np.array_equal(hv, U[2])
        out, g = sim.adjoint_step_grad(20, dts[-1], lam)
    assert np.array_equal(out, ref)
    assert g == pytest.approx(g_ref, rel=1e-12)


def test_resident_functional_equals_host_weights(oracle_lib):
    """dual_solver with the weights kept on the device (tg_set_functional,
    fn NULL; the first reverse step reads them in place) gives the sweep with
    host weights bit for bit, on a single member and on an ensemble, and a
    later sweep with host weights again the same."""
    pb = _pile_problem(128, 64, 0.3, 0.5, 25)
    _say("case resident functional")
    w = _box(pb, 0.2, 0.7, 0.2, 0.8)
    with Simulation(pb, checkpoint_every=6, use_graph=1) as sim:
        sim.run()
        lam_h, g_h = sim.dual_solver(w, want_dphi=True)
        sim.set_functional(w)
        lam_d, g_d = sim.dual_solver(None, want_dphi=True)
        lam_d2, g_d2 = sim.dual_solver(None, want_dphi=True)
        lam_h2, g_h2 = sim.dual_solver(w, want_dphi=True)  # and back to host weights
    assert np.array_equal(lam_h, lam_d) and np.array_equal(lam_h, lam_d2)
    assert np.array_equal(lam_h, lam_h2)
    assert g_h == g_d == g_d2 == g_h2
    mem = _members(12)
    weights = [_box(p, 0.2, 0.8, 0.2, 0.8).reshape(3, -1) for p in mem]
    with Ensemble(mem, precision=_lib.TG_FP64, checkpoint_every=5) as ens:
        ens.run(12)
        lam_h, g_h = ens.dual_solver(weights, want_dphi=True)
        ens.set_functional(weights)
        lam_d, g_d = ens.dual_solver(None, want_dphi=True)
        lam_h2, g_h2 = ens.dual_solver(weights, want_dphi=True)
    for a, b, c in zip(lam_h, lam_d, lam_h2):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    assert np.array_equal(g_h, g_d) and np.array_equal(g_h, g_h2)
