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
