"""GPU: the error indicator (tg_set_error_grid + tg_error_indicator) against
the oracle, bit for bit: the reconstructed previous state, state and
adjoint, the residual, the error and the correction of every error
element, on a small adaptive flow (one material and two), on a uniform
block, and on the 0.88 M-leaf adaptive mesh (3.5 M error elements: the
reconstruction checked on a sample of elements, the residual, error and
correction on all of them from the same reconstruction)."""
import sys

import numpy as np
import pytest

from discrete_adjoint_amd import errorgrid, problems
from discrete_adjoint_amd.solver import Simulation
from helpers import small_flow_problem, with_materials
import oracle
from oracle import error_oracle as eo
from test_error_grid import _leaf_case, _table_cells, error_oracle_of

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(got, rec, res, err, corr):
    assert np.array_equal(got["recon"], rec)
    assert np.array_equal(got["residual"], res)
    assert np.array_equal(got["error"], err)
    assert np.array_equal(got["correction"], corr)


@pytest.mark.parametrize("materials", [False, True])
def test_error_indicator_small_adaptive(oracle_lib, materials):
    pb, prev, state, lam, s, dt, eg, geom, cells = _leaf_case()
    if materials:
        pb = with_materials(pb, (16.0, 24.0))
        eg = errorgrid.build(geom, cells, lambda xc: (np.repeat(pb.grav[:, :1], len(xc), 1),
                                                      np.repeat(pb.curv[:, :1], len(xc), 1)),
                             mat=pb.mat)
    _say(f"case error grid small adaptive {pb.n} leaves, {eg.n} elements, materials={materials}")
    rec, res, err, corr, _ = error_oracle_of(pb, eg, geom, cells, prev, state, lam, s, dt)
    with Simulation(pb, max_steps=4) as sim:
        sim.set_error_grid(eg)
        got = sim.error_indicator(prev, state, lam, s, dt, want_recon=True)
        again = sim.error_indicator(prev, state, lam, s, dt, want_recon=True)
    _check(got, rec, res, err, corr)
    _check(again, rec, res, err, corr)
    assert np.abs(res).max() > 0 and err.max() > 0
    assert set(np.unique(eg.kind)) == {0, 1, 2, 3, 4}


def test_error_indicator_uniform_block(oracle_lib):
    pb = small_flow_problem(24, 20, 6)
    orc = oracle_lib.Oracle.from_problem(pb)
    U = [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]
    for s in range(4):
        prev = np.array([x.copy() for x in U])
        dt, _ = orc.step(U, s)
    state = np.array(U)
    rng = np.random.default_rng(2)
    lam = rng.normal(size=(3, pb.n))
    geom = (pb.nx, pb.ny, pb.x0 - 0.5 * pb.dx, pb.y0 - 0.5 * pb.dy, pb.dx, pb.dy)
    cells = errorgrid.uniform_cells(pb.nx, pb.ny)
    g = pb.grav[:, :1]
    eg = errorgrid.build(geom, cells, lambda xc: (np.repeat(g, len(xc), 1),
                                                  np.repeat(pb.curv[:, :1], len(xc), 1)))
    _say(f"case error grid uniform {pb.nx}x{pb.ny}, {eg.n} elements")
    rec, res, err, corr, _ = error_oracle_of(pb, eg, geom, cells, prev, state, lam, 3, dt)
    with Simulation(pb, max_steps=4) as sim:
        sim.set_error_grid(eg)
        got = sim.error_indicator(prev, state, lam, 3, dt, want_recon=True)
    _check(got, rec, res, err, corr)


def test_error_indicator_0p88m_leaves(oracle_lib):
    pb = problems.adaptive_fractal_problem(nsteps=4, checkpoint_every=4)
    geom = (pb.nx0, pb.ny0, pb.xmin, pb.ymin, pb.dx0, pb.dy0)
    cells = errorgrid.cells_of_table(pb.table, pb.xmin, pb.ymin)
    statics = problems.leaf_dem_statics(pb)
    eg = errorgrid.build(geom, cells, statics)
    _say(f"case error grid adaptive {pb.n} leaves, {eg.n} elements")
    with Simulation(pb, max_steps=4) as sim:
        sim.run(3)
        prev = np.array(sim.trajectory_state(2))
        state = np.array(sim.trajectory_state(3))
        dt = float(sim.dt_record()[2])
        w = pb.box_weights()
        lam, _ = sim.dual_solver(w)
        lam = np.array(lam).reshape(3, -1)
        sim.set_error_grid(eg)
        got = sim.error_indicator(prev, state, lam, 2, dt, want_recon=True)
        ms = sim.error_kernel_ms()
    _say("error kernels ms: " + ", ".join(f"{v:.3f}" for v in ms))
    # reconstruction on a sample of elements (the oracle's stencils)
    g = eo.Grid(geom, cells)
    tc = _table_cells(eg, geom)
    rng = np.random.default_rng(7)
    sample = rng.choice(eg.n, 3000, replace=False)
    fwd = np.concatenate([prev, state, lam])
    rec_s, _ = eo.reconstruct(g, [tc[q] for q in sample], fwd)
    assert np.array_equal(got["recon"][:, sample], rec_s)
    # residual, error, correction of every element from that reconstruction
    t = eg.table
    lo = oracle.LeafOracle(t.nbr, t.nkind, t.gen, t.dx, t.xc, eg.grav, eg.dgrav, eg.curv,
                           pb.phi_int_deg, pb.phi_bed_deg, pb.epsilon, pb.gravity_scale)
    res, err, corr = eo.indicator(lo, got["recon"], lam[:, eg.father], 2, dt)
    assert np.array_equal(got["residual"], res)
    assert np.array_equal(got["error"], err)
    assert np.array_equal(got["correction"], corr)
    assert err.max() > 0
