"""Adaptive-mesh flattening (host): Hilbert keys, hash buckets, traversal
order and the 8-slot neighbour tables.

Pins: the reference's own fhsfc2d_ (golden vectors in tests/golden, and the
reference source compiled into oracle/_ref when it is present), curve
properties on 2^k grids, the addElement list order restated independently,
and a geometric brute-force neighbour search.
"""
import json
import os

import numpy as np
import pytest

from discrete_adjoint_amd import mesh
from oracle import flatten_oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def test_keys_match_reference_golden_vectors():
    with open(os.path.join(HERE, "golden", "hsfc_keys.json")) as f:
        g = json.load(f)
    xy = np.array(g["norm_xy"])
    want = np.array(g["keys"], dtype=np.uint32)
    got = mesh.hsfc2d(xy)
    assert np.array_equal(got, want)


@pytest.mark.skipif(not os.path.exists(flatten_oracle.REF_HSFC),
                    reason="reference Hilbert source not built (make -C oracle ref)")
def test_keys_match_reference_build_random():
    rng = np.random.default_rng(7)
    xy = np.concatenate([rng.random((5000, 2)), rng.random((500, 2)) * 1.4 - 0.2])
    assert np.array_equal(mesh.hsfc2d(xy), flatten_oracle.ref_keys(xy))


@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_hilbert_curve_properties(k):
    # on a 2^k x 2^k grid of cell centres the leading 2k key bits enumerate
    # every cell once, consecutive cells are side neighbours, the curve starts
    # in the (0, 0) cell and ends in the (0, 2^k - 1) cell
    m = 1 << k
    xy = flatten_oracle.hilbert_cells(k)
    keys = mesh.hsfc2d(xy).astype(np.uint64)
    full = (keys[:, 0] << np.uint64(32)) | keys[:, 1]
    idx = (full >> np.uint64(64 - 2 * k)).astype(np.int64)
    assert sorted(idx.tolist()) == list(range(m * m))
    cell = np.empty((m * m, 2), dtype=np.int64)
    cell[idx] = np.floor(xy * m).astype(np.int64)
    steps = np.abs(np.diff(cell, axis=0)).sum(axis=1)
    assert (steps == 1).all()
    assert tuple(cell[0]) == (0, 0) and tuple(cell[-1]) == (0, m - 1)


def test_hash_bucket_formula():
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 2**32, size=(2000, 2), dtype=np.uint64).astype(np.uint32)
    kr = mesh.KEY_RANGE
    got = mesh.hash_bucket(keys, 100000)
    want = [flatten_oracle.bucket_of(k, kr, kr, 100000) for k in keys]
    assert got.tolist() == want
    # the bucket is monotone in the 64-bit key (8 wraps of the table)
    assert got.min() >= 0 and got.max() < 100000


@pytest.mark.parametrize("seed,nb", [(1, 100000), (2, 7), (3, 1), (4, 13)])
def test_flatten_matches_bruteforce(seed, nb):
    nx0, ny0 = 5, 4
    lev, ix, iy = mesh.random_balanced_tree(nx0, ny0, 3, 0.35, seed)
    t = mesh.flatten(nx0, ny0, 100.0, -50.0, 8.0, 6.0, lev, ix, iy, nbuckets=nb)
    # keys of the exact dyadic centres, normalised by the domain extent
    n = len(lev)
    cx = (ix + 0.5) / (nx0 * 2.0 ** lev)
    cy = (iy + 0.5) / (ny0 * 2.0 ** lev)
    keys = mesh.hsfc2d(np.stack([cx, cy], axis=1))
    ref = flatten_oracle.flatten(nx0, ny0, 100.0, -50.0, 8.0, 6.0, lev, ix, iy, nb, keys)
    assert np.array_equal(t.order, ref["order"])
    assert np.array_equal(t.key, keys[t.order])
    assert np.array_equal(t.nbr, ref["nbr"])
    assert np.array_equal(t.nkind, ref["nkind"])
    assert np.array_equal(t.ngen, ref["ngen"])
    assert np.array_equal(t.gen, lev[t.order])
    assert np.allclose(t.dx[:, 0], 8.0 / 2.0 ** lev[t.order], rtol=0, atol=0)
    assert np.array_equal(t.xc[:, 0], 100.0 + (ix[t.order] + 0.5) * (8.0 / 2.0 ** lev[t.order]))
    # with one bucket the traversal is the addElement order: sorted by key
    if nb == 1:
        k64 = (t.key[:, 0].astype(np.uint64) << np.uint64(32)) | t.key[:, 1]
        assert (np.diff(k64.astype(np.float64)) >= 0).all()
    assert sorted(t.order.tolist()) == list(range(n))


def test_bucket_list_quirk_equal_first_word():
    # addElement (hashtab2.C:185-217) compares the second word only once it
    # stands on an entry with an equal first word, and then walks on without
    # re-checking the first word: (5, 9) added to [(5, 1), (6, 0)] lands
    # after (6, 0).  Hand-worked from the loop; the oracle and the product
    # both give it.
    keys = np.array([[5, 1], [6, 0], [5, 9]], dtype=np.uint32)
    lst = []
    for e in range(3):
        flatten_oracle.add_element(lst, keys, e)
    assert lst == [0, 1, 2]
    assert mesh.traversal(keys, nbuckets=1).tolist() == [0, 1, 2]
    # insertion order matters once first words repeat
    assert mesh.traversal(keys[[2, 0, 1]], nbuckets=1).tolist() == [1, 0, 2]
    rng = np.random.default_rng(9)
    for nb in (1, 3, 100000):
        k = np.stack([rng.integers(0, 4, 300), rng.integers(0, 6, 300)], axis=1)
        k = (k * 1000003).astype(np.uint32)
        assert mesh.traversal(k, nbuckets=nb).tolist() == flatten_oracle.traversal(k, nb)


def test_neighbour_symmetry_and_levels():
    lev, ix, iy = mesh.random_balanced_tree(4, 4, 3, 0.4, 5)
    t = mesh.flatten(4, 4, 0.0, 0.0, 1.0, 1.0, lev, ix, iy)
    n = len(lev)
    for q in range(n):
        for s in range(8):
            f = t.nbr[q, s]
            if f < 0:
                continue
            assert abs(int(t.gen[f]) - int(t.gen[q])) <= 1
            # the neighbour lists q on the opposite side
            o = (s % 4 + 2) % 4
            assert q in (t.nbr[f, o], t.nbr[f, o + 4])


@pytest.mark.parametrize("case,code", [
    ("overlap", "overlapping"), ("gap", "cover"), ("unbalanced", "level")])
def test_flatten_errors(case, code):
    if case == "overlap":
        lev, ix, iy = [0, 0, 1], [0, 1, 0], [0, 0, 0]
        nx0 = 2
    elif case == "gap":
        lev, ix, iy = [0], [0], [0]
        nx0 = 2
    else:  # a level-2 leaf next to a level-0 leaf
        nx0 = 2
        lev, ix, iy = [0], [0], [0]
        for j in range(4):
            for i in range(4):
                lev.append(2)
                ix.append(4 + i)
                iy.append(j)
        lev, ix, iy = lev, ix, iy
    with pytest.raises(ValueError, match=code):
        mesh.flatten(nx0, 1, 0.0, 0.0, 1.0, 1.0, lev, ix, iy)
