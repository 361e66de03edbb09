"""TEST INFRASTRUCTURE ONLY: brute-force restatement of the reference's
adaptive-mesh traversal inputs, used to check tg_host_flatten.

- element keys: the reference's own fhsfc2d_ (src/main/hilbert.C:267-278)
  when oracle/_ref/libhsfc_ref.so has been built from the reference sources
  (`make -C oracle ref`), otherwise the committed golden vectors;
- hash bucket: HashTable::hash (hashtab.h:181-189) with hashconstant from
  the constructor (hashtab2.C:85), in Python floats (IEEE FP64, as the C++);
- bucket lists: HashTable::addElement (hashtab2.C:185-217) restated as a
  linked-list walk;
- neighbours: geometric search over leaf rectangles (every leaf whose
  rectangle shares a side segment of positive length), ordered by the corner
  rule of the son construction (refine2.C:589-700).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_HSFC = os.path.join(HERE, "_ref", "libhsfc_ref.so")


def ref_keys(norm_xy):
    """Keys from the reference's own fhsfc2d_ (None when _ref is not built)."""
    if not os.path.exists(REF_HSFC):
        return None
    lib = C.CDLL(REF_HSFC)
    lib.fhsfc2d_.restype = None
    lib.fhsfc2d_.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_uint), C.POINTER(C.c_uint)]
    xy = np.ascontiguousarray(norm_xy, dtype=np.float64).reshape(-1, 2)
    out = np.zeros((len(xy), 2), dtype=np.uint32)
    nkey = C.c_uint(2)
    key = (C.c_uint * 2)()
    for i, (x, y) in enumerate(xy):
        c = (C.c_double * 2)(x, y)
        lib.fhsfc2d_(c, C.byref(nkey), key)
        out[i] = (key[0], key[1])
    return out


def bucket_of(key, keyrange0, keyrange1, nbuckets):
    hc = 8.0 * nbuckets / (keyrange0 * keyrange1 + keyrange1)
    return int((float(key[0]) * keyrange1 + float(key[1])) * hc + 0.5) % nbuckets


def add_element(lst, keys, e):
    """Inserts e into a bucket list (list of leaf ids) as addElement does:
    walk while key0 is larger; on an equal key0, walk while key1 is larger."""
    k0, k1 = int(keys[e][0]), int(keys[e][1])
    p = 0
    while p < len(lst) and k0 > int(keys[lst[p]][0]):
        p += 1
    if p < len(lst) and k0 == int(keys[lst[p]][0]):
        while p < len(lst) and k1 > int(keys[lst[p]][1]):
            p += 1
    lst.insert(p, e)


def traversal(keys, nbuckets, keyrange=4294967296.0 + 1.0):
    lists = [[] for _ in range(nbuckets)]
    for e in range(len(keys)):
        add_element(lists[bucket_of(keys[e], keyrange, keyrange, nbuckets)], keys, e)
    return [e for lst in lists for e in lst]


def leaf_rects(nx0, ny0, xmin, ymin, dx0, dy0, lev, ix, iy):
    """Rectangles in integer units of the finest level present (exact)."""
    top = int(max(lev))
    r = []
    for L, i, j in zip(lev, ix, iy):
        s = 1 << (top - L)
        r.append((i * s, j * s, (i + 1) * s, (j + 1) * s))
    return r, top


def neighbours(nx0, ny0, lev, ix, iy):
    """Per leaf and side s (0 S, 1 E, 2 N, 3 W): the list of neighbouring
    leaves sharing a segment of that side, ordered from corner s to corner
    s+1 (corners 0 SW, 1 SE, 2 NE, 3 NW); [] on the domain boundary."""
    rects, top = leaf_rects(nx0, ny0, 0, 0, 1, 1, lev, ix, iy)
    W, H = nx0 << top, ny0 << top
    n = len(rects)
    out = []
    for e in range(n):
        x0, y0, x1, y1 = rects[e]
        sides = []
        for s in range(4):
            if (s == 0 and y0 == 0) or (s == 1 and x1 == W) or (s == 2 and y1 == H) or \
                    (s == 3 and x0 == 0):
                sides.append(None)
                continue
            found = []
            for f in range(n):
                if f == e:
                    continue
                a0, b0, a1, b1 = rects[f]
                if s == 0 and b1 == y0 and min(x1, a1) > max(x0, a0):
                    found.append((a0, f))
                elif s == 1 and a0 == x1 and min(y1, b1) > max(y0, b0):
                    found.append((b0, f))
                elif s == 2 and b0 == y1 and min(x1, a1) > max(x0, a0):
                    found.append((-a0, f))
                elif s == 3 and a1 == x0 and min(y1, b1) > max(y0, b0):
                    found.append((-b0, f))
            found.sort()
            sides.append([f for _, f in found])
        out.append(sides)
    return out


def flatten(nx0, ny0, xmin, ymin, dx0, dy0, lev, ix, iy, nbuckets, keys):
    """Reference-order tables from given leaf keys (input order = insertion
    order).  Returns dict of numpy arrays in flattened order."""
    order = traversal(keys, nbuckets)
    pos = {e: p for p, e in enumerate(order)}
    nb = neighbours(nx0, ny0, lev, ix, iy)
    n = len(lev)
    nbr = np.full((n, 8), -9, dtype=np.int64)
    nkind = np.full((n, 8), -9, dtype=np.int64)
    ngen = np.full((n, 8), -9, dtype=np.int64)
    for e in range(n):
        q = pos[e]
        for s in range(4):
            lst = nb[e][s]
            if lst is None:
                nbr[q, s] = nbr[q, s + 4] = -1
                nkind[q, s], nkind[q, s + 4] = -1, -2
                ngen[q, s] = ngen[q, s + 4] = lev[e]
            elif len(lst) == 1:
                nbr[q, s] = nbr[q, s + 4] = pos[lst[0]]
                nkind[q, s], nkind[q, s + 4] = 0, -2
                ngen[q, s] = ngen[q, s + 4] = lev[lst[0]]
            else:
                assert len(lst) == 2, "not 2:1 balanced"
                nbr[q, s], nbr[q, s + 4] = pos[lst[0]], pos[lst[1]]
                nkind[q, s] = nkind[q, s + 4] = 0
                ngen[q, s], ngen[q, s + 4] = lev[lst[0]], lev[lst[1]]
    return {"order": np.array(order), "nbr": nbr, "nkind": nkind, "ngen": ngen}


def hilbert_cells(k):
    """Top 2k bits of the key of every cell centre of a 2^k x 2^k grid, via the
    reference (if built) -- used by the curve-property checks."""
    m = 1 << k
    c = (np.arange(m) + 0.5) / m
    X, Y = np.meshgrid(c, c, indexing="ij")
    return np.stack([X.ravel(), Y.ravel()], axis=1)
