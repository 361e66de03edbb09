"""GPU: tile variants (shape and XCD-banded order) change the launch, not the
result: states, dt records and adjoints are bit-identical across variants;
only the order of the FP64 gradient partial sums moves (last bits)."""
import numpy as np
import pytest

from discrete_adjoint_amd import _lib
from discrete_adjoint_amd.solver import Ensemble
from helpers import small_flow_problem

pytestmark = pytest.mark.gpu


def _members():
    a = small_flow_problem(nx=40, ny=20, nsteps=25)
    b = small_flow_problem(nx=70, ny=33, nsteps=25, slope=20.0, phi_bed=18.0)
    return [a, b]


def _box(pb):
    L = pb.nx * pb.dx
    return pb.box_weights((0.45 * L, 0.8 * L, 0.0, pb.ny * pb.dy)).reshape(3, -1)


def _run(precision, **cfg):
    mem = _members()
    with Ensemble(mem, precision=precision, checkpoint_every=6, **cfg) as ens:
        ens.run(25)
        st = [ens.member_state(b) for b in range(2)]
        dts = [ens.member_dt_record(b) for b in range(2)]
        lam, dphi = ens.dual_solver([_box(p) for p in mem], want_dphi=True)
        sel = ens.tile_selection()
    return st, dts, lam, dphi, sel


@pytest.mark.parametrize("precision", [_lib.TG_FP64, _lib.TG_MIXED])
def test_tile_shapes_give_identical_results(precision):
    print(f"case tile shapes precision {precision}", flush=True)
    ref = _run(precision)
    assert ref[4] == [0, 0, 0]
    for tx, ty in ((64, 4), (32, 8), (16, 16)):
        got = _run(precision, tile_x=tx, tile_y=ty)
        for b in range(2):
            for x, y in zip(got[0][b], ref[0][b]):
                assert np.array_equal(x, y), (tx, ty)
            assert np.array_equal(got[1][b], ref[1][b])
            assert np.array_equal(got[2][b], ref[2][b])
        assert got[3] == pytest.approx(ref[3], rel=1e-12, abs=1e-300)


def test_autotune_selects_valid_variants_and_keeps_results():
    print("case autotune", flush=True)
    mem = _members()
    ref = _run(_lib.TG_FP64)
    with Ensemble(mem, precision=_lib.TG_FP64, checkpoint_every=6) as ens:
        ens.run(25)
        var, ms = ens.autotune(reps=3)
        assert all(0 <= v < _lib.TG_NUM_TILE_VARIANTS for v in var)
        assert ens.tile_selection() == var
        assert np.all(ms > 0)
        for w in range(3):
            assert ms[w, var[w]] == ms[w].min()
        # a fresh run with the tuned variants reproduces the default one
        h0 = [np.concatenate([getattr(p, f) for p in mem]) for f in ("h", "hu", "hv")]
        ens.set_state(*h0)
        ens.run(25)
        for b in range(2):
            for x, y in zip(ens.member_state(b), ref[0][b]):
                assert np.array_equal(x, y)
        lam, dphi = ens.dual_solver([_box(p) for p in mem], want_dphi=True)
        for b in range(2):
            assert np.array_equal(lam[b], ref[2][b])
        assert dphi == pytest.approx(ref[3], rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("precision,tiles", [(_lib.TG_FP64, (0, 0)), (_lib.TG_MIXED, (0, 0)),
                                             (_lib.TG_FP64, (16, 16)),
                                             (_lib.TG_FWD64_ADJ32, (32, 8))])
def test_graph_replay_equals_eager(precision, tiles):
    # the captured sweeps are replayed several times with eager work (and
    # synchronisations) in between; every replay must equal the eager run
    print(f"case graph replay precision {precision} tiles {tiles}", flush=True)
    ref = _run(precision)
    mem = _members()
    h0 = [np.concatenate([getattr(p, f) for p in mem]) for f in ("h", "hu", "hv")]
    with Ensemble(mem, precision=precision, checkpoint_every=6, use_graph=1,
                  tile_x=tiles[0], tile_y=tiles[1]) as ens:
        for rep in range(3):
            ens.set_state(*h0)
            ens.run(25)
            for b in range(2):
                for x, y in zip(ens.member_state(b), ref[0][b]):
                    assert np.array_equal(x, y), rep
                assert np.array_equal(ens.member_dt_record(b), ref[1][b])
            lam, dphi = ens.dual_solver([_box(p) for p in mem], want_dphi=True)
            for b in range(2):
                assert np.array_equal(lam[b], ref[2][b]), rep
            assert dphi == pytest.approx(ref[3], rel=1e-12, abs=1e-300)
            # eager work between replays: the next set_state uploads the state
            # and clears the error flag on the same stream, then synchronises
            ens.get_state()


@pytest.mark.parametrize("frac,graph", [(1.0, 0), (0.4, 0), (0.4, 1), (0.04, 1)])
def test_stored_trajectory_equals_checkpointing(frac, graph):
    # keeping the last states of the run whole (no recompute for them) gives
    # the same adjoint and gradient bit for bit
    print(f"case stored trajectory fraction {frac} graph {graph}", flush=True)
    ref = _run(_lib.TG_FP64)
    mem = _members()
    state = 8 * 3 * sum(p.n for p in mem)
    with Ensemble(mem, precision=_lib.TG_FP64, checkpoint_every=6, use_graph=graph,
                  store_bytes=int(frac * 25 * state)) as ens:
        for rep in range(2):
            h0 = [np.concatenate([getattr(p, f) for p in mem]) for f in ("h", "hu", "hv")]
            ens.set_state(*h0)
            ens.run(25)
            for b in range(2):
                for x, y in zip(ens.member_state(b), ref[0][b]):
                    assert np.array_equal(x, y)
            lam, dphi = ens.dual_solver([_box(p) for p in mem], want_dphi=True)
            for b in range(2):
                assert np.array_equal(lam[b], ref[2][b]), (frac, rep)
            assert np.array_equal(np.asarray(dphi), np.asarray(ref[3]))
