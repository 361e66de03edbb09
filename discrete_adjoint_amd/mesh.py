"""Adaptive-mesh flattening: numpy views of tg_host_hsfc2d,
tg_host_hash_bucket and tg_host_flatten (include/tg_host.h), and the
trajectory spill files (tg_host_write_solution / tg_host_read_solution)."""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib

# element table size of the reference (datread.C:560: EL_TABLE_SIZE)
EL_TABLE_SIZE = 100000
KEY_RANGE = 4294967296.0 + 1.0


def hsfc2d(norm_xy):
    """Hilbert keys (n, 2) uint32 of normalised coordinates (n, 2)."""
    xy = np.ascontiguousarray(norm_xy, dtype=np.float64).reshape(-1, 2)
    keys = np.zeros((len(xy), 2), dtype=np.uint32)
    rc = _lib.load().tg_host_hsfc2d(len(xy), _lib.dptr(xy.reshape(-1)), _lib.uptr(keys))
    if rc:
        raise ValueError(f"tg_host_hsfc2d returned {rc}")
    return keys


def leaf_d_gravity(table, gz):
    """calc_d_gravity over the leaves of a flattened table: (2, n)."""
    n = len(table.gen)
    nbr = np.ascontiguousarray(table.nbr, dtype=np.int32).reshape(-1)
    xc = np.ascontiguousarray(table.xc, dtype=np.float64).reshape(-1)
    g = np.ascontiguousarray(gz, dtype=np.float64).reshape(-1)
    out = np.zeros(2 * n)
    rc = _lib.load().tg_host_leaf_d_gravity(n, _lib.iptr(nbr), _lib.dptr(xc), _lib.dptr(g),
                                            _lib.dptr(out))
    if rc:
        raise ValueError(f"tg_host_leaf_d_gravity returned {rc}")
    return out.reshape(2, n)


def write_solution(path, keys, U, order=None):
    """One step's trajectory spill file in the SolRec layout
    (dualmesh.C:87-142): keys (n, 2) uint32, U = (h, hu, hv); elements in
    `order` (None: index order)."""
    k = np.ascontiguousarray(keys, dtype=np.uint32).reshape(-1, 2)
    h, hu, hv = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in U)
    o = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
    rc = _lib.load().tg_host_write_solution(str(path).encode(), len(k), _lib.uptr(k),
                                            None if o is None else _lib.iptr(o),
                                            _lib.dptr(h), _lib.dptr(hu), _lib.dptr(hv))
    if rc:
        raise OSError(f"tg_host_write_solution returned {rc}")


def read_solution(path, keys, U=None):
    """Reads a spill file into the elements with `keys` (dualmesh.C:234-297).
    U: initial (h, hu, hv) for elements absent from the file (None: NaN).
    Returns ((h, hu, hv), found, unknown)."""
    k = np.ascontiguousarray(keys, dtype=np.uint32).reshape(-1, 2)
    n = len(k)
    h, hu, hv = ([np.full(n, np.nan) for _ in range(3)] if U is None else
                 [np.array(a, dtype=np.float64, copy=True).reshape(-1) for a in U])
    found, unknown = C.c_int64(), C.c_int64()
    rc = _lib.load().tg_host_read_solution(str(path).encode(), n, _lib.uptr(k), _lib.dptr(h),
                                           _lib.dptr(hu), _lib.dptr(hv), C.byref(found),
                                           C.byref(unknown))
    if rc:
        raise OSError(f"tg_host_read_solution returned {rc}")
    return (h, hu, hv), found.value, unknown.value


def hash_bucket(keys, nbuckets=EL_TABLE_SIZE, keyrange=(KEY_RANGE, KEY_RANGE)):
    k = np.ascontiguousarray(keys, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros(len(k), dtype=np.int32)
    rc = _lib.load().tg_host_hash_bucket(len(k), _lib.uptr(k), keyrange[0], keyrange[1],
                                         nbuckets, _lib.iptr(out))
    if rc:
        raise ValueError(f"tg_host_hash_bucket returned {rc}")
    return out


def traversal(keys, nbuckets=EL_TABLE_SIZE, keyrange=(KEY_RANGE, KEY_RANGE)):
    """Element-table traversal order of keys added in index order."""
    k = np.ascontiguousarray(keys, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros(len(k), dtype=np.int32)
    rc = _lib.load().tg_host_traversal(len(k), _lib.uptr(k), keyrange[0], keyrange[1],
                                       nbuckets, _lib.iptr(out))
    if rc:
        raise ValueError(f"tg_host_traversal returned {rc}")
    return out


@dataclass
class LeafTable:
    order: np.ndarray   # (n,) flattened position -> input leaf index
    key: np.ndarray     # (n, 2) uint32
    gen: np.ndarray     # (n,)
    nbr: np.ndarray     # (n, 8) flattened neighbour index, -1 boundary
    nkind: np.ndarray   # (n, 8) 0 present, -1 boundary, -2 no second neighbour
    ngen: np.ndarray    # (n, 8)
    dx: np.ndarray      # (n, 2)
    xc: np.ndarray      # (n, 2)


FLATTEN_ERRORS = {-1: "bad argument", -2: "overlapping or out-of-range leaves",
                  -3: "leaves do not cover the base grid",
                  -4: "neighbours more than one level apart"}


def flatten(nx0, ny0, xmin, ymin, dx0, dy0, lev, ix, iy, nbuckets=EL_TABLE_SIZE):
    lev, ix, iy = (np.ascontiguousarray(a, dtype=np.int32) for a in (lev, ix, iy))
    n = len(lev)
    t = LeafTable(order=np.zeros(n, np.int32), key=np.zeros((n, 2), np.uint32),
                  gen=np.zeros(n, np.int32), nbr=np.zeros((n, 8), np.int32),
                  nkind=np.zeros((n, 8), np.int32), ngen=np.zeros((n, 8), np.int32),
                  dx=np.zeros((n, 2)), xc=np.zeros((n, 2)))
    c = _lib.tg_leaf_table()
    c.order, c.key, c.gen = _lib.iptr(t.order), _lib.uptr(t.key), _lib.iptr(t.gen)
    c.nbr, c.nkind, c.ngen = _lib.iptr(t.nbr), _lib.iptr(t.nkind), _lib.iptr(t.ngen)
    c.dx, c.xc = _lib.dptr(t.dx), _lib.dptr(t.xc)
    rc = _lib.load().tg_host_flatten(nx0, ny0, xmin, ymin, dx0, dy0, n, _lib.iptr(lev),
                                     _lib.iptr(ix), _lib.iptr(iy), nbuckets, C.byref(c))
    if rc:
        raise ValueError(f"tg_host_flatten: {FLATTEN_ERRORS.get(rc, rc)}")
    return t


def random_balanced_tree(nx0, ny0, max_level, frac, seed):
    """A seeded 2:1-balanced quadtree over an nx0 x ny0 base grid: leaves are
    refined at random (probability frac per leaf and level) and neighbours
    are refined until no two side-adjacent leaves differ by more than one
    level.  Returns (lev, ix, iy): base cells row by row, each refined cell
    replaced in place by its sons 0..3 (SW, SE, NE, NW)."""
    rng = np.random.default_rng(seed)
    leaves = [(0, i, j) for j in range(ny0) for i in range(nx0)]

    def split(c):
        L, i, j = c
        return [(L + 1, 2 * i, 2 * j), (L + 1, 2 * i + 1, 2 * j), (L + 1, 2 * i + 1, 2 * j + 1),
                (L + 1, 2 * i, 2 * j + 1)]

    for L in range(max_level):
        out = []
        for c in leaves:
            out.extend(split(c) if c[0] == L and rng.random() < frac else [c])
        leaves = out
    # balance: refine any leaf with a neighbour two or more levels finer
    while True:
        S = set(leaves)

        def finest_at(L, i, j, depth):
            # is any leaf strictly inside cell (L, i, j) at level >= L + depth?
            for d in range(depth, max_level - L + 1):
                s = 1 << d
                for a in range(i * s, (i + 1) * s):
                    for b in range(j * s, (j + 1) * s):
                        if (L + d, a, b) in S:
                            return True
            return False

        todo = set()
        for (L, i, j) in leaves:
            W, H = nx0 << L, ny0 << L
            for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                a, b = i + di, j + dj
                if 0 <= a < W and 0 <= b < H and finest_at(L, a, b, 2):
                    todo.add((L, i, j))
        if not todo:
            break
        out = []
        for c in leaves:
            out.extend(split(c) if c in todo else [c])
        leaves = out
    lev = np.array([c[0] for c in leaves], np.int32)
    ix = np.array([c[1] for c in leaves], np.int32)
    iy = np.array([c[2] for c in leaves], np.int32)
    return lev, ix, iy


def _dilate4(a):
    d = a.copy()
    d[1:, :] |= a[:-1, :]
    d[:-1, :] |= a[1:, :]
    d[:, 1:] |= a[:, :-1]
    d[:, :-1] |= a[:, 1:]
    return d


def _up(a):
    return np.repeat(np.repeat(a, 2, axis=0), 2, axis=1)


def refine_tree(nx0, ny0, max_level, want):
    """A 2:1-balanced quadtree over an nx0 x ny0 base grid.  want(L, I, J)
    returns, for the level-L grid of cell indices (I, J), which cells to split
    (consulted only where the cell exists).  Splits are then added from fine
    to coarse until side-adjacent leaves differ by at most one level: a split
    level-(L+1) cell needs its four side neighbours to exist, i.e. their
    level-L parents split.  Returns (lev, ix, iy), coarse leaves first, row by
    row within a level."""
    R = []
    for L in range(max_level):
        J, I = np.mgrid[0:(ny0 << L), 0:(nx0 << L)]
        r = np.asarray(want(L, I, J), dtype=bool).copy()
        if L > 0:
            r &= _up(R[L - 1])
        R.append(r)
    for L in range(max_level - 1, 0, -1):
        d = _dilate4(R[L])
        H, W = d.shape
        R[L - 1] |= d.reshape(H // 2, 2, W // 2, 2).any(axis=(1, 3))
    lev, ix, iy = [], [], []
    for L in range(max_level + 1):
        exists = np.ones((ny0, nx0), bool) if L == 0 else _up(R[L - 1])
        leaf = exists & ~R[L] if L < max_level else exists
        jj, ii = np.nonzero(leaf)
        lev.append(np.full(len(ii), L, np.int32))
        ix.append(ii.astype(np.int32))
        iy.append(jj.astype(np.int32))
    return np.concatenate(lev), np.concatenate(ix), np.concatenate(iy)


# ---- mesh adaptation (tg_host_adapt, tg_host_adapt_forward / _adjoint) ----------------

@dataclass
class AdaptPass:
    """One adaptation pass: the leaves before and after (tree order: lev, ix,
    iy arrays) and the split / merge operations in the order applied
    (structured array kind, lev, ix, iy)."""
    old: tuple
    new: tuple
    ops: np.ndarray

    def forward(self, fields):
        """The state transfer applied to fields on the old leaves (nf, n_old)
        -> (nf, n_new)."""
        f = np.ascontiguousarray(fields, dtype=np.float64)
        nf = f.shape[0]
        out = np.empty((nf, len(self.new[0])))
        o, nw = [np.ascontiguousarray(a, dtype=np.int32) for a in self.old], \
            [np.ascontiguousarray(a, dtype=np.int32) for a in self.new]
        rc = _lib.load().tg_host_adapt_forward(
            len(self.ops), _ops_ptr(self.ops), nf, len(o[0]), *[_lib.iptr(a) for a in o],
            _lib.dptr(f.reshape(-1)), len(nw[0]), *[_lib.iptr(a) for a in nw],
            _lib.dptr(out.reshape(-1)))
        if rc:
            raise ValueError(f"tg_host_adapt_forward returned {rc}")
        return out

    def adjoint(self, lam_new):
        """The transposed transfer: adjoint (3, n_new) on the new leaves ->
        (3, n_old) on the old ones."""
        lam = np.ascontiguousarray(lam_new, dtype=np.float64).reshape(3, -1)
        out = np.empty((3, len(self.old[0])))
        o, nw = [np.ascontiguousarray(a, dtype=np.int32) for a in self.old], \
            [np.ascontiguousarray(a, dtype=np.int32) for a in self.new]
        rc = _lib.load().tg_host_adapt_adjoint(
            len(self.ops), _ops_ptr(self.ops), len(nw[0]), *[_lib.iptr(a) for a in nw],
            _lib.dptr(lam.reshape(-1)), len(o[0]), *[_lib.iptr(a) for a in o],
            _lib.dptr(out.reshape(-1)))
        if rc:
            raise ValueError(f"tg_host_adapt_adjoint returned {rc}")
        return out


OP_DTYPE = np.dtype([("kind", np.int32), ("lev", np.int32), ("ix", np.int32), ("iy", np.int32)])


def _ops_ptr(ops):
    if len(ops) == 0:
        return None
    return ops.ctypes.data_as(C.POINTER(_lib.tg_adapt_op))


def adapt(nx0, ny0, lev, ix, iy, U, max_level, num_buffer_layer=5, tiny=1e-4, el_error=None):
    """One adaptation pass of the tree (lev, ix, iy) with state U = (h, hu,
    hv) on its leaves: returns (lev, ix, iy, U) of the new tree and the
    AdaptPass."""
    lib = _lib.load()
    lv, x, y = [np.ascontiguousarray(a, dtype=np.int32) for a in (lev, ix, iy)]
    h, hu, hv = [np.ascontiguousarray(a, dtype=np.float64) for a in U]
    err = None if el_error is None else np.ascontiguousarray(el_error, dtype=np.float64)
    prm = _lib.tg_adapt_params(max_level, num_buffer_layer, tiny)
    n = len(lv)
    cap, op_cap = 4 * n + 64, 4 * n + 64
    while True:
        out = [np.empty(cap, dtype=np.int32) for _ in range(3)]
        Uo = [np.empty(cap) for _ in range(3)]
        ops = np.zeros(op_cap, dtype=OP_DTYPE)
        n_out, nops = C.c_int(), C.c_int()
        rc = lib.tg_host_adapt(nx0, ny0, C.byref(prm), n, _lib.iptr(lv), _lib.iptr(x),
                               _lib.iptr(y), _lib.dptr(h), _lib.dptr(hu), _lib.dptr(hv),
                               None if err is None else _lib.dptr(err), cap, C.byref(n_out),
                               *[_lib.iptr(a) for a in out], *[_lib.dptr(a) for a in Uo],
                               op_cap, C.byref(nops), _ops_ptr(ops))
        if rc == -2:
            cap, op_cap = max(cap, n_out.value), max(op_cap, nops.value)
            continue
        if rc:
            raise ValueError(f"tg_host_adapt returned {rc}")
        m = n_out.value
        new = tuple(a[:m].copy() for a in out)
        return (*new, tuple(a[:m].copy() for a in Uo),
                AdaptPass((lv, x, y), new, ops[:nops.value].copy()))
