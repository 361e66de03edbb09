"""CPU checks of the C-ABI library: it loads, exports every function
include/*.h declares, and its host-side set-up helpers (terrain coefficients,
piles, functional weights) reproduce closed forms.  No GPU compute here."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from discrete_adjoint_amd import _lib, problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(tg_\w+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared_functions()
    assert "tg_forward_run" in declared and "tg_adjoint_run" in declared
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers them all
    assert declared <= set(_lib.SIGNATURES), declared - set(_lib.SIGNATURES)


def test_config_defaults_are_the_reference_constants():
    lib = _lib.load()
    cfg = _lib.tg_config()
    lib.tg_config_default(C.byref(cfg))
    assert cfg.tiny == 1e-4            # constant.h:68
    assert cfg.gamma == 1.0            # hpfem.C:82
    assert cfg.cfl_inner == 0.5 and cfg.cfl_outer == 0.5
    assert cfg.frict_tiny_early == 0.1 and cfg.frict_tiny_late == 1e-9
    assert cfg.frict_tiny_switch == 50


def test_params_from_angles_closed_form():
    lib = _lib.load()
    p = _lib.tg_params()
    assert lib.tg_params_from_angles(math.radians(30), math.radians(20), 1.0, 1.0,
                                     C.byref(p)) == 0
    c, t = math.cos(math.radians(30)), math.tan(math.radians(20))
    assert p.sin_int == math.sin(math.radians(30))
    assert p.tan_bed == t
    assert p.kact_c1 == 2.0 / (c * c)
    assert p.kact_root == math.sqrt(abs(1.0 - (1.0 + t * t) * c * c))
    # d root / d phi_bed by central FD
    def root(phb):
        tt = math.tan(phb)
        return math.sqrt(abs(1.0 - (1.0 + tt * tt) * c * c))
    e = 1e-6
    fd = (root(math.radians(20) + e) - root(math.radians(20) - e)) / (2 * e)
    assert p.dkroot_dphi == pytest.approx(fd, rel=1e-7)


def test_no_device_is_reported_not_crashed():
    # On a machine without a HIP device tg_create must fail cleanly.
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    lib = _lib.load()
    ctx = C.c_void_p()
    cfg = _lib.tg_config()
    lib.tg_config_default(C.byref(cfg))
    rc = lib.tg_create(C.byref(cfg), C.byref(ctx))
    assert rc in (_lib.TG_ERR_NODEVICE, _lib.TG_ERR_HIP)


def test_tile_variants_are_described_and_validated():
    lib = _lib.load()
    shapes = []
    for v in range(_lib.TG_NUM_TILE_VARIANTS):
        tx, ty, band = C.c_int(), C.c_int(), C.c_int()
        assert lib.tg_tile_variant(v, C.byref(tx), C.byref(ty), C.byref(band)) == 0
        assert tx.value * ty.value == 256
        shapes.append((tx.value, ty.value, band.value))
    assert shapes[0] == (64, 4, 0)       # the plain row-major table
    assert {(a, b) for a, b, c in shapes if c} == {(64, 4), (32, 8), (16, 16)}
    assert lib.tg_tile_variant(_lib.TG_NUM_TILE_VARIANTS, None, None, None) == _lib.TG_ERR_ARG
    # an unbuilt tile shape is refused before any device is touched
    cfg = _lib.tg_config()
    lib.tg_config_default(C.byref(cfg))
    cfg.tile_x, cfg.tile_y = 128, 2
    ctx = C.c_void_p()
    assert lib.tg_create(C.byref(cfg), C.byref(ctx)) == _lib.TG_ERR_UNSUPPORTED


def test_horn_slope_is_exact_on_a_plane():
    # z = a x + b y sampled on the raster: Horn's operator and the bilinear
    # sampling return (a, b); gravity is the plane's; curvature and
    # d gravity vanish (GisApi.C:1295-1345, element2.C:3545-3653)
    a, b = -0.25, 0.125
    res, nr, nc = 5.0, 60, 80
    xmin, ymax = 0.0, (nr - 1) * res
    cols = np.arange(nc) * res + xmin
    rows = ymax - np.arange(nr) * res
    X, Y = np.meshgrid(cols, rows)
    elev = np.ascontiguousarray((a * X + b * Y + 100.0).astype(np.float32).reshape(-1))
    nx, ny, dx = 30, 20, 10.0
    g, dg, cv = problems.static_coefs((nr, nc, xmin, ymax, res, elev), nx, ny, dx, dx,
                                      25.0, 25.0)
    ms = math.hypot(a, b)
    ang = math.atan(ms)
    np.testing.assert_allclose(g[0], -9.8 * math.sin(ang) * a / ms, rtol=2e-6)
    np.testing.assert_allclose(g[1], -9.8 * math.sin(ang) * b / ms, rtol=2e-6)
    np.testing.assert_allclose(g[2], 9.8 * math.cos(ang), rtol=2e-6)
    assert np.abs(cv).max() < 1e-6
    assert np.abs(dg).max() < 1e-6


def test_flat_dem_gives_vertical_gravity():
    nr = nc = 40
    elev = np.full(nr * nc, 12.0, dtype=np.float32)
    g, dg, cv = problems.static_coefs((nr, nc, 0.0, 195.0, 5.0, elev), 10, 10, 10.0, 10.0,
                                      50.0, 50.0)
    assert np.all(g[0] == 0) and np.all(g[1] == 0) and np.all(g[2] == 9.8)
    assert not dg.any() and not cv.any()


def test_pile_volume_converges_to_paraboloid():
    # elliptical paraboloid volume pi/2 * a * b * H; the 9-node average is a
    # second-order quadrature of it
    for n, tol in ((64, 2e-2), (256, 2e-3)):
        L = 1000.0
        dx = L / n
        h, hu, hv = problems.piles(n, n, dx, dx, 0.5 * dx, 0.5 * dx,
                                   [(500.0, 480.0, 300.0, 200.0, 30.0, 20.0)])
        vol = h.sum() * dx * dx
        assert vol == pytest.approx(math.pi / 2 * 300 * 200 * 20, rel=tol)
        assert not hu.any() and not hv.any()
        assert h.min() >= 0


def test_box_weights():
    w = problems.Problem("t", 10, 10, 10.0, 10.0, 5.0, 5.0, None, None, None, None, None,
                         None, 30, 20, 1, 1).box_weights((20.0, 60.0, 0.0, 30.0))
    w = w.reshape(3, 10, 10)
    assert not w[1:].any()
    # centres 25,35,45,55 in x; 5,15,25 in y
    assert np.count_nonzero(w[0]) == 4 * 3 and np.all(w[0][w[0] > 0] == 100.0)


def test_fractal_dem_is_seeded_and_scaled():
    a = problems.fractal_dem(129, 800.0, 0.8, 7)
    b = problems.fractal_dem(129, 800.0, 0.8, 7)
    c = problems.fractal_dem(129, 800.0, 0.8, 8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() == 0.0 and a.max() == pytest.approx(800.0, rel=1e-6)
