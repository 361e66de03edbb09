"""TEST INFRASTRUCTURE ONLY: CPU restatement of the reference's error
indicator on the error grid, the checker of tg_host_error_stencil and
tg_error_indicator.  Imported by tests/ only.

- The error grid: every cell of the mesh refined once (uinform_refine,
  uniform_refine.C:24-130); sons 0..3 = SW, SE, NE, NW of their father.
- Stencils (bilinear_interp, bilinear_interp.C:126-289): the twelve side
  neighbours of a family of sons (Element::get_side_neighbor,
  element2.C:4224-4330), here found by a geometric search over the sons'
  (level, i, j) cells, and the coincidence cases; the points are the
  fathers' centres computed from each element's centre and which_son
  (Neighbor, :55-106), merged within 1e-12 (:439-459), two points sorted by
  NeighLess (:114-121).
- Reconstruction (bilinear_interp_elem, :427-556): one point copies the
  father's values, two interpolate linearly as linear_interp is written
  (:346-349, no constant term), three barycentrically (:398-408), four by the
  4x4 fit a0 + a1 x + a2 y + a3 x y solved with dgesv (:350-383), restated
  below in the order of the reference LAPACK (dgetrf's partial pivoting,
  reciprocal column scaling and rank-one updates; dgetrs' interchanges and
  two triangular solves).
- Residual, error and correction (update_bilinear_error_grid,
  error_grid_routines.C:13-53; error_compute, error_compute.C:24-100): the
  leaf oracle's orl_error_update on the error grid, then res = predicted -
  reconstructed state, error = sum |res_j (lambda_j - lambda~_j)|,
  correction = sum res_j lambda_j with lambda the father's adjoint.
"""
import numpy as np

SONS = ((0, 0), (1, 0), (1, 1), (0, 1))
DI, DJ = (0, 1, 0, -1), (-1, 0, 1, 0)          # sides 0 S, 1 E, 2 N, 3 W
# get_side_neighbor's sides 0..3 relative to positive_x_side = 1: x+, y+, x-, y-
SIDE_OF = {0: 1, 1: 2, 2: 3, 3: 0}
COPY, LINEAR_X, LINEAR_Y, BARYCENTRIC, BILINEAR = range(5)


class Grid:
    """The error grid of a mesh given by its cells (lev, ix, iy) over an
    nx0 x ny0 base grid of dx0 x dy0 cells from (xmin, ymin)."""

    def __init__(self, geometry, cells):
        self.nx0, self.ny0, self.xmin, self.ymin, self.dx0, self.dy0 = geometry
        lev, ix, iy = (np.asarray(a).tolist() for a in cells)
        self.el = {}      # (L, i, j) -> (father, which_son)
        for f, (L, i, j) in enumerate(zip(lev, ix, iy)):
            for k, (a, b) in enumerate(SONS):
                self.el[(L + 1, 2 * i + a, 2 * j + b)] = (f, k)

    def size(self, c):
        return self.dx0 / (1 << c[0]), self.dy0 / (1 << c[0])

    def centre(self, c):
        hx, hy = self.size(c)
        return self.xmin + (c[1] + 0.5) * hx, self.ymin + (c[2] + 0.5) * hy

    def _inside(self, L, i, j):
        return 0 <= i < (self.nx0 << L) and 0 <= j < (self.ny0 << L)

    def side(self, c, s):
        """Neighbours across side s: [] at the domain boundary, [one] or the
        two finer ones [next to corner s, next to corner s + 1]."""
        L, i, j = c
        i2, j2 = i + DI[s], j + DJ[s]
        if not self._inside(L, i2, j2):
            return []
        if (L, i2, j2) in self.el:
            return [(L, i2, j2)]
        if (L - 1, i2 >> 1, j2 >> 1) in self.el:
            return [(L - 1, i2 >> 1, j2 >> 1)]
        ci, cj = 2 * i2, 2 * j2
        kids = {0: [(ci, cj + 1), (ci + 1, cj + 1)], 1: [(ci, cj), (ci, cj + 1)],
                2: [(ci + 1, cj), (ci, cj)], 3: [(ci + 1, cj + 1), (ci + 1, cj)]}[s]
        out = [(L + 1, a, b) for a, b in kids]
        assert all(o in self.el for o in out), "not 2:1 balanced"
        return out

    def get_side_neighbor(self, c, k):
        if c is None:
            return None
        if k < 4:
            nb = self.side(c, SIDE_OF[k])
            return nb[0] if nb else None
        if k < 8:   # the second neighbour of a side with two, else the one
            nb = self.side(c, SIDE_OF[k - 4])
            return nb[-1] if nb else None
        first, then = {8: (1, 0), 9: (2, 1), 10: (3, 2), 11: (0, 3)}[k]
        return self.get_side_neighbor(self.get_side_neighbor(c, first), then)

    def two_on(self, c, k):
        """neigh_proc[k + 4] >= 0: two neighbours on side k (0..3 as above)."""
        return len(self.side(c, SIDE_OF[k])) == 2

    def point(self, c):
        x, y = self.centre(c)
        hx, hy = self.size(c)
        w = self.el[c][1]
        sx, sy = {0: (1, 1), 1: (-1, 1), 2: (-1, -1), 3: (1, -1)}[w]
        return (x + sx * .5 * hx, y + sy * .5 * hy, self.el[c][0])

    def stencil(self, c):
        """(kind, [(x, y, father leaf)]) of error element c."""
        L, i, j = c
        f, w = self.el[c]
        base = (L, i - SONS[w][0], j - SONS[w][1])
        son = [(L, base[1] + a, base[2] + b) for a, b in SONS]
        g = self.get_side_neighbor
        nb = {3: g(son[0], 3), 6: g(son[0], 6), 10: g(son[0], 10), 0: g(son[1], 0),
              7: g(son[1], 7), 11: g(son[1], 11), 8: g(son[2], 8), 4: g(son[2], 4),
              1: g(son[2], 1), 5: g(son[3], 5), 9: g(son[3], 9), 2: g(son[3], 2)}

        def eq(a, b):
            return nb[a] is not None and nb[b] is not None and nb[a] == nb[b]
        e, XM, YM, XP, YP = c, 2, 3, 0, 1
        if w == 0:
            if eq(6, 10):
                st = (nb[10], nb[3], nb[2] if self.two_on(e, XM) else nb[9], e)
            elif eq(3, 10):
                st = (nb[10], nb[7] if self.two_on(e, YM) else nb[11], nb[6], e)
            else:
                st = (nb[10], nb[3], nb[6], e)
        elif w == 1:
            if eq(11, 0):
                st = (nb[7], nb[11], e, nb[4] if self.two_on(e, XP) else nb[8])
            elif eq(7, 11):
                st = (nb[3] if self.two_on(e, YM) else nb[10], nb[11], e, nb[0])
            else:
                st = (nb[7], nb[11], e, nb[0])
        elif w == 2:
            if eq(4, 8):
                st = (e, nb[0] if self.two_on(e, XP) else nb[11], nb[1], nb[8])
            elif eq(8, 1):
                st = (e, nb[4], nb[5] if self.two_on(e, YP) else nb[9], nb[8])
            else:
                st = (e, nb[4], nb[1], nb[8])
        else:
            if eq(9, 2):
                st = (nb[6] if self.two_on(e, XM) else nb[10], e, nb[9], nb[5])
            elif eq(9, 5):
                st = (nb[2], e, nb[9], nb[1] if self.two_on(e, YP) else nb[8])
            else:
                st = (nb[2], e, nb[9], nb[5])
        u = []
        for q in st:
            if q is None:
                continue
            p = self.point(q)
            if not any(abs(p[0] - v[0]) < 1e-12 and abs(p[1] - v[1]) < 1e-12 for v in u):
                u.append(p)
        if len(u) == 1:
            return COPY, u
        if len(u) == 2:
            a, b = u
            if (b[0] < a[0] and abs(a[0] - b[0]) > 1e-12) or (abs(a[0] - b[0]) < 1e-12
                                                              and b[1] < a[1]):
                u = [b, a]
            return (LINEAR_Y if u[1][0] - u[0][0] < 1e-12 else LINEAR_X), u
        return (BARYCENTRIC if len(u) == 3 else BILINEAR), u


def dgesv4(A, b):
    """Solve A x = b (4x4) in the arithmetic order of the reference LAPACK's
    dgesv; returns (x, singular).  With an exactly zero pivot dgesv does not
    solve and b comes back unchanged."""
    A = [list(map(float, r)) for r in A]
    b = list(map(float, b))
    piv = []
    singular = False
    for k in range(4):
        p = k
        for i in range(k + 1, 4):
            if abs(A[i][k]) > abs(A[p][k]):
                p = i
        piv.append(p)
        if A[p][k] != 0.0:
            if p != k:
                A[k], A[p] = A[p], A[k]
            r = 1.0 / A[k][k]
            for i in range(k + 1, 4):
                A[i][k] = A[i][k] * r
        else:
            singular = True
        for j in range(k + 1, 4):
            t = -A[k][j]
            if t != 0.0:
                for i in range(k + 1, 4):
                    A[i][j] = A[i][j] + A[i][k] * t
    if singular:
        return b, True
    for i in range(4):
        b[i], b[piv[i]] = b[piv[i]], b[i]
    for k in range(4):
        if b[k] != 0.0:
            for i in range(k + 1, 4):
                b[i] = b[i] - b[k] * A[i][k]
    for k in range(3, -1, -1):
        if b[k] != 0.0:
            b[k] = b[k] / A[k][k]
            for i in range(k):
                b[i] = b[i] - b[k] * A[i][k]
    return b, False


def interpolate(kind, u, f, x, y):
    """bilinear_interp_elem's value at (x, y) from points u and values f."""
    if kind == COPY:
        return f[0]
    if kind == LINEAR_X:
        return (f[1] - f[0]) / (u[1][0] - u[0][0]) * (x - u[0][0])
    if kind == LINEAR_Y:
        return (f[1] - f[0]) / (u[1][1] - u[0][1]) * (y - u[0][1])
    if kind == BARYCENTRIC:
        (x1, y1, _), (x2, y2, _), (x3, y3, _) = u
        inv_det = 1 / ((y2 - y3) * (x1 - x3) + (x3 - x2) * (y1 - y3))
        l1 = ((y2 - y3) * (x - x3) + (x3 - x2) * (y - y3)) * inv_det
        l2 = ((y3 - y1) * (x - x3) + (x1 - x3) * (y - y3)) * inv_det
        l3 = 1 - l1 - l2
        return l1 * f[0] + l2 * f[1] + l3 * f[2]
    if f[0] == 0.0 and f[1] == 0.0 and f[2] == 0.0 and f[3] == 0.0:
        return 0.0
    A = [[1.0, p[0], p[1], p[0] * p[1]] for p in u]
    c, _ = dgesv4(A, f)
    return c[0] + c[1] * x + c[2] * y + c[3] * x * y


def reconstruct(grid, table_cells, fwd):
    """Reconstructed fields (9, n) of the error elements listed in
    table_cells ((lev, ix, iy) in the order to return) from fwd (9, nf):
    previous state, state, adjoint of the mesh cells."""
    out = np.empty((9, len(table_cells)))
    stencils = []
    for q, c in enumerate(table_cells):
        kind, u = grid.stencil(c)
        stencils.append((kind, u))
        x, y = grid.centre(c)
        f0 = grid.el[c][0]
        for v in range(9):
            vals = [fwd[v][f0]] if kind == COPY else [fwd[v][p[2]] for p in u]
            out[v, q] = interpolate(kind, u, vals, x, y)
    return out, stencils


def indicator(leaf_oracle, recon, lam_father, s, dt):
    """Residual (3, n), error (n,), correction (n,) from the reconstruction
    (9, n) and the fathers' adjoint (3, n), on the error grid's leaf oracle."""
    pred = leaf_oracle.error_update(recon[0:3], s, dt)
    res = np.empty((3, recon.shape[1]))
    err = np.zeros(recon.shape[1])
    corr = np.zeros(recon.shape[1])
    for j in range(3):
        res[j] = pred[j] - recon[3 + j]
        err += np.abs(res[j] * (lam_father[j] - recon[6 + j]))
        corr += res[j] * lam_father[j]
    return res, err, corr
