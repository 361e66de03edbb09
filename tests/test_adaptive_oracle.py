"""The adaptive time loop's reverse sweep, on the CPU restatement
(oracle/adapt_oracle.py): lambda^0 and dJ/dphi_bed of a run that re-meshes
every 4 steps against central finite differences of the whole run at frozen
dt and frozen mesh operations -- the adjoint crosses every pass through the
transposed transfer (sum of the sons' adjoints for a split, 0.25 of the
father's for a merge).  A planted mistake in the transposed transfer (0.5
instead of 0.25) fails the same check."""
import math

import numpy as np
import pytest

from helpers import box_weights_of, small_adaptive_run
from oracle import adapt_oracle

BOX = (90.0, 170.0, 50.0, 150.0)


def _J(run, res):
    return float(np.sum(box_weights_of(run, res["tree"], BOX) * res["U"]))


def test_adaptive_reverse_sweep_matches_fd(oracle_lib):
    run = small_adaptive_run()
    N = 14
    w = lambda tree: box_weights_of(run, tree, BOX)
    ref = adapt_oracle.run(run, N, w)
    kinds = [None if ops is None else set(ops["kind"].tolist()) for _, ops in ref["passes"]]
    assert any(k and 0 in k for k in kinds) and any(k and 1 in k for k in kinds)
    frozen = lambda k, tree, U: ref["passes"][k]
    lam0 = ref["lam0"]
    rng = np.random.default_rng(3)
    wet = np.nonzero(run.U[0] > 0.5)[0]
    for e in rng.choice(wet, 4, replace=False):
        vals = []
        eps = 1e-6 * run.U[0][e]
        for sgn in (1, -1):
            U0 = np.array(run.U, copy=True)
            U0[0][e] += sgn * eps
            r = adapt_oracle.run(run, N, w, ops_of=frozen, U0=U0, dt_seq=ref["dts"],
                                 want_dphi=False)
            vals.append(_J(run, r))
        fd = (vals[0] - vals[1]) / (2 * eps)
        assert lam0[0][e] == pytest.approx(fd, rel=1e-5, abs=1e-7), e
    # dJ/dphi_bed through the passes
    vals = []
    for sgn in (1, -1):
        r = adapt_oracle.run(run, N, w, ops_of=frozen, dt_seq=ref["dts"], want_dphi=False,
                             phi_bed_deg=run.phi_bed_deg + sgn * 1e-4)
        vals.append(_J(run, r))
    fd = (vals[0] - vals[1]) / (2 * math.radians(1e-4))
    assert ref["dphi"] != 0 and ref["dphi"] == pytest.approx(fd, rel=2e-4)


def test_planted_transfer_mistake_is_caught(oracle_lib, monkeypatch):
    run = small_adaptive_run()
    N = 14
    w = lambda tree: box_weights_of(run, tree, BOX)
    ref = adapt_oracle.run(run, N, w, want_dphi=False)
    frozen = lambda k, tree, U: ref["passes"][k]
    orig = adapt_oracle.transfer_adjoint

    def wrong(ops, new_tree, lam_new, old_tree):
        out = orig(ops, new_tree, lam_new, old_tree)
        merged = [o for o in ops if o["kind"] == 1]
        return out * (2.0 if merged else 1.0)   # a son taking 0.5, not 0.25, of its father
    monkeypatch.setattr(adapt_oracle, "transfer_adjoint", wrong)
    bad = adapt_oracle.run(run, N, w, ops_of=frozen, dt_seq=ref["dts"], want_dphi=False)
    monkeypatch.setattr(adapt_oracle, "transfer_adjoint", orig)
    e = int(np.argmax(run.U[0]))
    eps = 1e-6 * run.U[0][e]
    vals = []
    for sgn in (1, -1):
        U0 = np.array(run.U, copy=True)
        U0[0][e] += sgn * eps
        vals.append(_J(run, adapt_oracle.run(run, N, w, ops_of=frozen, U0=U0, dt_seq=ref["dts"],
                                             want_dphi=False)))
    fd = (vals[0] - vals[1]) / (2 * eps)
    assert ref["lam0"][0][e] == pytest.approx(fd, rel=1e-5)
    assert not (bad["lam0"][0][e] == pytest.approx(fd, rel=1e-3))


def test_oracle_transfer_equals_the_product_transfer():
    from discrete_adjoint_amd import mesh
    run = small_adaptive_run()
    U = np.array(run.U)
    lev, ix, iy, Un, p = mesh.adapt(*run.geometry[:2], *run.tree, U, run.max_level, 2)
    assert len(p.ops)
    new = (lev, ix, iy)
    assert np.array_equal(adapt_oracle.transfer_forward(p.ops, run.tree, U, new), np.array(Un))
    lam = np.random.default_rng(5).normal(size=(3, len(lev)))
    assert np.array_equal(adapt_oracle.transfer_adjoint(p.ops, new, lam, run.tree),
                          p.adjoint(lam))
