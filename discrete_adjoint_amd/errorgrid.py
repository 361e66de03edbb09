"""The error indicator on the error grid (the reference's Error build:
dual_solver.C:143-175, 208-240).

The error grid is the mesh with every cell refined once (uinform_refine,
uniform_refine.C:24-130).  Its elements carry their father's state,
previous state and adjoint; bilinear_interp (bilinear_interp.C:126-556)
reconstructs all three from a stencil of neighbouring elements, and
error_compute (error_compute.C:24-100) weighs the residual of one step of the
reconstructed previous state with the adjoint.

build() makes the error grid of a mesh on the host (the sons' tree,
tg_host_flatten's tables, tg_host_error_stencil's stencils, the sons' static
coefficients); Simulation.set_error_grid hands it to the device and
Simulation.error_indicator evaluates it (tg_error_indicator).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib, mesh

SONS = ((0, 0), (1, 0), (1, 1), (0, 1))   # which_son 0..3: SW, SE, NE, NW

INTERP_COPY, INTERP_LINEAR_X, INTERP_LINEAR_Y, INTERP_BARYCENTRIC, INTERP_BILINEAR = range(5)


@dataclass
class ErrorGrid:
    table: object          # mesh.LeafTable of the error elements (flattened order)
    father: np.ndarray     # (n,) cell of the forward mesh each element refines
    which_son: np.ndarray  # (n,)
    kind: np.ndarray       # (n,) interpolation rule
    pts: np.ndarray        # (n, 4) forward cells of the stencil points (-1 unused)
    pxy: np.ndarray        # (n, 4, 2) their coordinates
    grav: np.ndarray       # (3, n)
    dgrav: np.ndarray      # (2, n)
    curv: np.ndarray       # (2, n)
    mat: np.ndarray = None

    @property
    def n(self):
        return len(self.father)

    def c_struct(self):
        """The tg_error_grid of these arrays (kept alive by this object)."""
        t = self.table
        keep = [np.ascontiguousarray(t.nbr, dtype=np.int32).reshape(-1),
                np.ascontiguousarray(t.nkind, dtype=np.int32).reshape(-1),
                np.ascontiguousarray(t.gen, dtype=np.int32),
                np.ascontiguousarray(t.dx, dtype=np.float64).reshape(-1),
                np.ascontiguousarray(t.xc, dtype=np.float64).reshape(-1),
                np.ascontiguousarray(self.grav, dtype=np.float64).reshape(-1),
                np.ascontiguousarray(self.dgrav, dtype=np.float64).reshape(-1),
                np.ascontiguousarray(self.curv, dtype=np.float64).reshape(-1),
                np.ascontiguousarray(self.father, dtype=np.int32),
                np.ascontiguousarray(self.kind, dtype=np.int32),
                np.ascontiguousarray(self.pts, dtype=np.int32).reshape(-1),
                np.ascontiguousarray(self.pxy, dtype=np.float64).reshape(-1)]
        g = _lib.tg_error_grid()
        g.n = self.n
        g.nbr, g.nproc, g.gen = (_lib.iptr(a) for a in keep[:3])
        g.cdx, g.xc, g.grav, g.dgrav, g.curv = (_lib.dptr(a) for a in keep[3:8])
        g.father, g.kind, g.pts = (_lib.iptr(a) for a in keep[8:11])
        g.pxy = _lib.dptr(keep[11])
        if self.mat is not None:
            m = np.ascontiguousarray(self.mat, dtype=np.uint8)
            keep.append(m)
            g.mat = m.ctypes.data_as(C.POINTER(C.c_uint8))
        self._keep = keep
        return g


def cells_of_table(table, xmin, ymin):
    """(lev, ix, iy) of every cell of a flattened leaf table."""
    lev = np.asarray(table.gen, dtype=np.int32)
    ix = np.rint((table.xc[:, 0] - xmin) / table.dx[:, 0] - 0.5).astype(np.int32)
    iy = np.rint((table.xc[:, 1] - ymin) / table.dx[:, 1] - 0.5).astype(np.int32)
    return lev, ix, iy


def stencil(table, which_son, father, nfather):
    """tg_host_error_stencil on a flattened error grid: (kind (n,), pts (n, 4),
    pxy (n, 4, 2))."""
    n = len(father)
    nbr = np.ascontiguousarray(table.nbr, dtype=np.int32)
    nk = np.ascontiguousarray(table.nkind, dtype=np.int32)
    dx = np.ascontiguousarray(table.dx, dtype=np.float64)
    xc = np.ascontiguousarray(table.xc, dtype=np.float64)
    ws = np.ascontiguousarray(which_son, dtype=np.int32)
    fa = np.ascontiguousarray(father, dtype=np.int32)
    kind = np.empty(n, np.int32)
    pts = np.empty((n, 4), np.int32)
    pxy = np.empty((n, 4, 2))
    rc = _lib.load().tg_host_error_stencil(n, _lib.iptr(nbr), _lib.iptr(nk), _lib.dptr(dx),
                                           _lib.dptr(xc), _lib.iptr(ws), _lib.iptr(fa), nfather,
                                           _lib.iptr(kind), _lib.iptr(pts), _lib.dptr(pxy))
    if rc:
        raise ValueError(f"tg_host_error_stencil returned {rc}")
    return kind, pts, pxy


def build(geometry, cells, statics, mat=None):
    """The error grid of a mesh.  geometry: (nx0, ny0, xmin, ymin, dx0, dy0)
    of the base grid; cells: (lev, ix, iy) of the mesh's cells in the order
    the context holds them (a uniform block: level 0 row by row; a leaf
    mesh: cells_of_table of its table); statics(xc (n, 2)) -> (grav (3, n),
    curv (2, n)) at the sons' centres (as the reference sets up every new
    son, calc_topo_data / calc_gravity_vector); mat: the cells' materials
    (the sons inherit them)."""
    nx0, ny0, xmin, ymin, dx0, dy0 = geometry
    lev, ix, iy = (np.asarray(a, dtype=np.int64) for a in cells)
    nf = len(lev)
    a = np.array([s[0] for s in SONS])
    b = np.array([s[1] for s in SONS])
    slev = np.repeat(lev + 1, 4).astype(np.int32)
    six = (2 * ix[:, None] + a[None, :]).reshape(-1).astype(np.int32)
    siy = (2 * iy[:, None] + b[None, :]).reshape(-1).astype(np.int32)
    sfather = np.repeat(np.arange(nf, dtype=np.int32), 4)
    swhich = np.tile(np.arange(4, dtype=np.int32), nf)
    t = mesh.flatten(nx0, ny0, xmin, ymin, dx0, dy0, slev, six, siy)
    o = t.order
    father, which = sfather[o], swhich[o]
    kind, pts, pxy = stencil(t, which, father, nf)
    grav, curv = statics(t.xc)
    grav = np.asarray(grav, dtype=np.float64).reshape(3, -1)
    curv = np.asarray(curv, dtype=np.float64).reshape(2, -1)
    dgrav = mesh.leaf_d_gravity(t, grav[2])
    m = None if mat is None else np.asarray(mat, dtype=np.uint8)[father]
    return ErrorGrid(t, father, which, kind, pts, pxy, grav, dgrav, curv, m)


def uniform_cells(nx, ny):
    """(lev, ix, iy) of a uniform nx x ny block, row by row."""
    j, i = np.divmod(np.arange(nx * ny), nx)
    return np.zeros(nx * ny, np.int32), i.astype(np.int32), j.astype(np.int32)
