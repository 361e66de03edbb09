"""Mesh adaptation between steps (tg_host_adapt: H_adapt then unrefine,
hadpt.C:50-323, unrefine.C:54-352) and its state / adjoint transfers.

Pins: a hand-worked pass (a pile in the middle cell of a 3 x 3 grid: the
cell and its four side neighbours split, nothing else; the second pass
changes nothing); untouched families merge with the 0.25-weighted state;
on random 2:1 trees with random piles: the result is a 2:1-balanced tiling
within max_level, every contour leaf below max_level is split, merged
families had no finer neighbour, mass (sum h * area) is conserved, the
state equals the forward transfer of the operations, and the adjoint
transfer is its transpose (dot-product identity); repeated passes reach a
fixed point with every contour leaf at max_level.
"""
import numpy as np
import pytest

from discrete_adjoint_amd import mesh


def _grid(nx, ny):
    return (np.zeros(nx * ny, int), np.tile(np.arange(nx), ny), np.repeat(np.arange(ny), nx))


def _cells(lev, ix, iy):
    return set(zip(np.asarray(lev).tolist(), np.asarray(ix).tolist(), np.asarray(iy).tolist()))


def test_hand_worked_pile_in_the_middle():
    lev, ix, iy = _grid(3, 3)
    h = np.zeros(9)
    h[4] = 1.0
    l2, x2, y2, U2, p = mesh.adapt(3, 3, lev, ix, iy, (h, 0 * h, 0 * h), max_level=1,
                                   num_buffer_layer=1)
    split = {(int(o["lev"]), int(o["ix"]), int(o["iy"])) for o in p.ops if o["kind"] == 0}
    assert split == {(0, 1, 1), (0, 1, 0), (0, 0, 1), (0, 2, 1), (0, 1, 2)}
    assert not (p.ops["kind"] == 1).any()
    assert len(l2) == 4 + 5 * 4
    # the four corner cells stay, sons copy their father
    assert {(0, 0, 0), (0, 2, 0), (0, 0, 2), (0, 2, 2)} <= _cells(l2, x2, y2)
    centre = (l2 == 1) & (x2 // 2 == 1) & (y2 // 2 == 1)
    assert centre.sum() == 4 and np.all(U2[0][centre] == 1.0) and U2[0].sum() == 4.0
    # sons follow their father's place, SW SE NE NW
    k = np.nonzero(centre)[0]
    assert list(zip(x2[k], y2[k])) == [(2, 2), (3, 2), (3, 3), (2, 3)]
    l3, x3, y3, U3, p3 = mesh.adapt(3, 3, l2, x2, y2, U2, max_level=1, num_buffer_layer=1)
    assert len(p3.ops) == 0 and _cells(l3, x3, y3) == _cells(l2, x2, y2)


def test_untouched_families_merge():
    # a 2 x 2 base grid refined once everywhere, no pile: every family merges
    lev = np.ones(16, int)
    ix = np.tile(np.arange(4), 4)
    iy = np.repeat(np.arange(4), 4)
    rng = np.random.default_rng(1)
    U = [rng.random(16) * 1e-6 for _ in range(3)]   # below every contour
    l2, x2, y2, U2, p = mesh.adapt(2, 2, lev, ix, iy, U, max_level=2)
    assert len(l2) == 4 and np.all(l2 == 0) and (p.ops["kind"] == 1).all()
    for e in range(4):
        sons = [np.nonzero((ix == 2 * x2[e] + a) & (iy == 2 * y2[e] + b))[0][0]
                for a, b in ((0, 0), (1, 0), (1, 1), (0, 1))]
        for v in range(3):
            want = 0.0
            for s in sons:
                want += U[v][s] * 0.25
            assert U2[v][e] == want


def _contour_leaves(nx0, ny0, lev, ix, iy, h, tiny):
    t = mesh.flatten(nx0, ny0, 0.0, 0.0, 1.0, 1.0, lev, ix, iy)
    hf = h[t.order]
    out = np.zeros(len(h), bool)
    for c in (tiny, 5 * tiny, 15 * tiny, 40 * tiny):
        for e in range(len(h)):
            nb = [t.nbr[e, k] for k in range(8) if t.nkind[e, k] == 0]
            if hf[e] >= c and any(hf[q] < c for q in nb):
                out[t.order[e]] = True
            if hf[e] < c and any(hf[q] >= c for q in nb):
                out[t.order[e]] = True
    return out


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_random_trees_invariants(seed):
    nx0, ny0, L = 6, 5, 3
    lev, ix, iy = mesh.random_balanced_tree(nx0, ny0, L - 1, 0.3, seed)
    rng = np.random.default_rng(seed)
    s = 1.0 / (1 << lev)
    X, Y = (ix + 0.5) * s, (iy + 0.5) * s
    cx, cy = rng.uniform(1, nx0 - 1), rng.uniform(1, ny0 - 1)
    h = np.maximum(0.0, 0.01 * (1.0 - ((X - cx) ** 2 + (Y - cy) ** 2) / 2.0))
    U = (h, rng.normal(size=len(h)) * h, rng.normal(size=len(h)) * h)
    l2, x2, y2, U2, p = mesh.adapt(nx0, ny0, lev, ix, iy, U, max_level=L)
    # a 2:1-balanced tiling (flatten checks coverage, overlap and balance)
    mesh.flatten(nx0, ny0, 0.0, 0.0, 1.0, 1.0, l2, x2, y2)
    assert l2.max() <= L
    # every contour leaf below max_level was split
    on = _contour_leaves(nx0, ny0, lev, ix, iy, h, 1e-4)
    new = _cells(l2, x2, y2)
    for e in np.nonzero(on & (lev < L))[0]:
        assert (lev[e], ix[e], iy[e]) not in new
    # merged families had no finer neighbour (and came from untouched leaves)
    t_old = mesh.flatten(nx0, ny0, 0.0, 0.0, 1.0, 1.0, lev, ix, iy)
    pos = {c: i for i, c in enumerate(zip(t_old.gen.tolist(), *(
        (np.floor(t_old.xc[:, d] / t_old.dx[:, d]).astype(int)).tolist() for d in (0, 1))))}
    for o in p.ops[p.ops["kind"] == 1]:
        for a, b in ((0, 0), (1, 0), (1, 1), (0, 1)):
            e = pos[(o["lev"] + 1, 2 * o["ix"] + a, 2 * o["iy"] + b)]
            assert not (t_old.nkind[e, 4:] == 0).any()
    # mass conserved, state = forward transfer of the operations
    area = lambda lv: 1.0 / 4.0 ** lv
    assert np.sum(U2[0] * area(l2)) == pytest.approx(np.sum(h * area(lev)), rel=1e-13)
    F = p.forward(np.array(U))
    for v in range(3):
        assert np.array_equal(F[v], U2[v])
    # the adjoint transfer is the transpose of the forward one
    x = rng.normal(size=(3, len(lev)))
    lam = rng.normal(size=(3, len(l2)))
    lhs = np.sum(lam * p.forward(x))
    rhs = np.sum(p.adjoint(lam) * x)
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_repeated_passes_reach_max_level_on_the_contours():
    nx0 = ny0 = 8
    L = 3
    lev, ix, iy = _grid(nx0, ny0)

    def pile(lev, ix, iy):
        s = 1.0 / (1 << lev)
        X, Y = (ix + 0.5) * s, (iy + 0.5) * s
        return np.maximum(0.0, 0.02 * (1.0 - ((X - 4.1) ** 2 + (Y - 3.7) ** 2) / 4.0))

    for it in range(L + 3):
        h = pile(lev, ix, iy)
        U = (h, 0 * h, 0 * h)
        lev, ix, iy, _, p = mesh.adapt(nx0, ny0, lev, ix, iy, U, max_level=L)
    h = pile(lev, ix, iy)
    _, _, _, _, p = mesh.adapt(nx0, ny0, lev, ix, iy, (h, 0 * h, 0 * h), max_level=L)
    assert len(p.ops) == 0          # a fixed point
    on = _contour_leaves(nx0, ny0, lev, ix, iy, h, 1e-4)
    assert on.any() and np.all(lev[on] == L)


def test_adjoint_transfer_hand_worked():
    # split (0, 0, 0) then merge its sons back: the adjoint of the pair
    ops = np.array([(0, 0, 0, 0), (1, 0, 0, 0)], dtype=mesh.OP_DTYPE)
    cell = (np.array([0]), np.array([0]), np.array([0]))
    p = mesh.AdaptPass(cell, cell, ops)
    lam = np.array([[4.0], [8.0], [-2.0]])
    # merge^T: each son 0.25 of the father; split^T: the father sums the sons
    assert np.array_equal(p.adjoint(lam), lam)
    one = mesh.AdaptPass(cell, (np.ones(4, int), np.array([0, 1, 1, 0]), np.array([0, 0, 1, 1])),
                         ops[:1])
    lam4 = np.array([[1.0, 2.0, 3.0, 4.0], [0.5, 0.5, 0.5, 0.5], [0, 0, 0, 1e-3]])
    assert np.array_equal(one.adjoint(lam4), [[10.0], [2.0], [1e-3]])
