"""Small synthetic problems for the tests (analytic coefficients, no DEM)."""
import math

import numpy as np

from discrete_adjoint_amd.problems import Problem


def analytic_problem(nx, ny, dx=10.0, dy=None, slope_x_deg=0.0, slope_y_deg=0.0,
                     h=None, hu=None, hv=None, phi_int=30.0, phi_bed=20.0, nsteps=10,
                     curv=(0.0, 0.0), checkpoint_every=50):
    """A uniform block on a plane (constant gravity vector, no curvature
    unless given).  h/hu/hv: arrays or callables f(X, Y) of centre coordinates."""
    dy = dx if dy is None else dy
    n = nx * ny
    x0, y0 = 0.5 * dx, 0.5 * dy
    X, Y = np.meshgrid(x0 + dx * np.arange(nx), y0 + dy * np.arange(ny))
    X, Y = X.reshape(-1), Y.reshape(-1)
    zx = -math.tan(math.radians(slope_x_deg))
    zy = -math.tan(math.radians(slope_y_deg))
    ms = math.sqrt(zx * zx + zy * zy)
    g = np.zeros((3, n))
    if ms > 0:
        ang = math.atan(ms)
        dsg = 9.8 * math.sin(ang)
        g[0] = -dsg * zx / ms
        g[1] = -dsg * zy / ms
        g[2] = 9.8 * math.cos(ang)
    else:
        g[2] = 9.8
    dg = np.zeros((2, n))
    cv = np.zeros((2, n))
    cv[0], cv[1] = curv

    def field(f):
        if f is None:
            return np.zeros(n)
        if callable(f):
            return np.ascontiguousarray(f(X, Y), dtype=np.float64)
        return np.ascontiguousarray(f, dtype=np.float64).reshape(-1)

    return Problem("analytic", nx, ny, dx, dy, x0, y0, g, dg, cv, field(h), field(hu),
                   field(hv), phi_int, phi_bed, nsteps, checkpoint_every)


def paraboloid(xc, yc, a, b, H):
    def f(X, Y):
        v = H * (1.0 - ((X - xc) / a) ** 2 - ((Y - yc) / b) ** 2)
        return np.maximum(v, 0.0)
    return f


def small_flow_problem(nx=24, ny=20, nsteps=30, slope=25.0, phi_bed=15.0):
    """A paraboloid pile released on an inclined plane (moving, wet/dry
    fronts, both friction terms active)."""
    L = nx * 10.0
    return analytic_problem(nx, ny, 10.0, slope_x_deg=slope, slope_y_deg=3.0,
                            h=paraboloid(0.35 * L, 0.5 * ny * 10.0, 60.0, 45.0, 8.0),
                            phi_int=32.0, phi_bed=phi_bed, nsteps=nsteps, curv=(1e-3, -5e-4))


def analytic_leaf_problem(nx0, ny0, dx0, max_level, want, slope_x_deg=0.0, slope_y_deg=0.0,
                          h=None, hu=None, hv=None, phi_int=30.0, phi_bed=20.0, nsteps=10,
                          curv=(0.0, 0.0), checkpoint_every=50, nbuckets=None):
    """An adaptive mesh (refine_tree(want), 2:1 balanced, flattened into the
    element-table order) on a plane: constant gravity vector as in
    analytic_problem; h/hu/hv callables f(X, Y) of the leaf centres."""
    from discrete_adjoint_amd import mesh
    from discrete_adjoint_amd.problems import LeafProblem
    lev, ix, iy = mesh.refine_tree(nx0, ny0, max_level, want)
    t = mesh.flatten(nx0, ny0, 0.0, 0.0, dx0, dx0, lev, ix, iy,
                     nbuckets=nbuckets or mesh.EL_TABLE_SIZE)
    n = len(lev)
    base = analytic_problem(1, 1, dx0, slope_x_deg=slope_x_deg, slope_y_deg=slope_y_deg,
                            curv=curv)
    g = np.repeat(base.grav, n, axis=1)
    dg = np.zeros((2, n))
    cv = np.repeat(base.curv, n, axis=1)
    X, Y = t.xc[:, 0], t.xc[:, 1]

    def field(f):
        return np.zeros(n) if f is None else np.ascontiguousarray(f(X, Y), dtype=np.float64)

    return LeafProblem("analytic_leaves", nx0, ny0, dx0, dx0, 0.0, 0.0, t, g, dg, cv, field(h),
                       field(hu), field(hv), phi_int, phi_bed, nsteps, checkpoint_every)


def disc(cx, cy, r):
    """want() of refine_tree: split cells whose centre is within r (+ one
    cell) of (cx, cy), r in base-cell units."""
    def want(L, I, J):
        s = 1.0 / (1 << L)
        return np.hypot((I + 0.5) * s - cx, (J + 0.5) * s - cy) < r + s
    return want


def small_leaf_flow_problem(nx0=16, ny0=12, max_level=2, nsteps=30, slope=25.0, phi_bed=15.0):
    """The small_flow_problem pile on an adaptive mesh refined around it:
    moving fronts cross hanging sides of both kinds."""
    dx0 = 20.0
    L = nx0 * dx0
    return analytic_leaf_problem(nx0, ny0, dx0, max_level, disc(0.35 * nx0, 0.5 * ny0, 3.0),
                                 slope_x_deg=slope, slope_y_deg=3.0,
                                 h=paraboloid(0.35 * L, 0.5 * ny0 * dx0, 60.0, 45.0, 8.0),
                                 phi_int=32.0, phi_bed=phi_bed, nsteps=nsteps,
                                 curv=(1e-3, -5e-4))


def with_materials(pb, phis, split=0.5):
    """pb with two bed-friction materials: material 0 (phis[0] deg) where the
    cell / leaf centre lies left of `split` of the x extent, material 1
    (phis[1] deg) right of it."""
    import copy
    q = copy.copy(pb)
    if hasattr(pb, "table"):
        x = pb.table.xc[:, 0]
        L = pb.nx0 * pb.dx0
    else:
        x = np.tile(pb.x0 + pb.dx * np.arange(pb.nx), pb.ny)
        L = pb.nx * pb.dx
    q.mat = (x >= split * L).astype(np.uint8)
    q.materials = tuple(phis)
    return q


def small_adaptive_run(nx0=12, ny0=10, max_level=2, adapt_every=4, num_buffer_layer=2,
                       slope=25.0, phi_bed=15.0):
    """An adaptive-run set-up (discrete_adjoint_amd.adaptive.AdaptiveRun) on a
    plane: the small_leaf_flow_problem pile on a tree refined around it,
    constant gravity and curvature at every leaf centre."""
    from discrete_adjoint_amd import adaptive, mesh
    dx0 = 20.0
    L = nx0 * dx0
    lev, ix, iy = mesh.refine_tree(nx0, ny0, max_level, disc(0.35 * nx0, 0.5 * ny0, 2.0))
    s = dx0 / (1 << lev)
    X, Y = (ix + 0.5) * s, (iy + 0.5) * s
    h = paraboloid(0.35 * L, 0.5 * ny0 * dx0, 50.0, 40.0, 6.0)(X, Y)
    U = np.stack([h, np.zeros_like(h), np.zeros_like(h)])
    base = analytic_problem(1, 1, dx0, slope_x_deg=slope, slope_y_deg=3.0, curv=(1e-3, -5e-4))

    def statics(xc):
        n = len(xc)
        return np.repeat(base.grav, n, axis=1), np.repeat(base.curv, n, axis=1)

    return adaptive.AdaptiveRun((nx0, ny0, 0.0, 0.0, dx0, dx0), (lev, ix, iy), U, statics, 32.0,
                                phi_bed, max_level, adapt_every=adapt_every,
                                num_buffer_layer=num_buffer_layer, checkpoint_every=3)


def box_weights_of(run, tree, box):
    """Weights (3, n) of 'h integrated over the box' on a tree (tree order)."""
    nx0, ny0, xmin, ymin, dx0, dy0 = run.geometry
    lev, ix, iy = tree
    s = 1.0 / (1 << np.asarray(lev))
    X, Y = xmin + (ix + 0.5) * dx0 * s, ymin + (iy + 0.5) * dy0 * s
    inside = (X > box[0]) & (X < box[1]) & (Y > box[2]) & (Y < box[3])
    w = np.zeros((3, len(lev)))
    w[0, inside] = (dx0 * s * dy0 * s)[inside]
    return w
