"""Closed-form pins of the oracle's per-cell rules that the end-to-end checks
(finite differences, symmetry, conservation) leave open.  Every expected value
below comes from the mathematics or from hand-worked numbers, never from the
oracle's own expressions:

- earth-pressure coefficient (getcoef.f:19-66): active / passive branch by the
  sign of the velocity divergence, the neutral value, the |u|,|v| < tiny reset,
  the clamp at 0, h <= tiny, the epsilon factor;
- slope limiter (element2.C:830-939, geoflow.h:90-100): literal stencils for
  the min of three, opposite signs, the c_sgn dead zone and its 1/2 factor;
- Coulomb bed friction of update_states (residual.C:162-275): a uniform sheet
  on a tilted plane accelerates at g (sin theta - cos theta tan phi_bed) (or
  + when moving upslope), curvature adds u^2 kappa to the normal load, the
  max(., 0) lift-off clamp, the frict_tiny dead zone, and no friction at rest;
- curvature of the terrain (GisApi.C:764-833 with calculate_slope
  GisApi.C:1295-1345): 2a, 2b on a quadratic DEM, the halved values and the
  extrapolated bands at the raster edges;
- element-table hash bucket (hashtab.h:181-189, hashtab2.C:85): exact rational
  buckets of keys within 1e-8 .. 2e-5 of a bucket boundary, in every one of
  the eight wraps of the table.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle
from discrete_adjoint_amd import mesh, problems


def _flat_oracle(nx, ny, dx, x0, y0, phi_int, phi_bed, eps=1.0, gx=0.0, gz=9.8, kx=0.0,
                 gamma=1.0):
    n = nx * ny
    grav = np.stack([np.full(n, gx), np.zeros(n), np.full(n, gz)])
    curv = np.stack([np.full(n, kx), np.zeros(n)])
    return oracle.Oracle(nx, ny, dx, dx, x0, y0, grav, np.zeros((2, n)), curv, phi_int,
                         phi_bed, epsilon=eps, gamma=gamma)


def _earth_pressure(phi_int_deg, phi_bed_deg):
    """Savage-Hutter active / passive coefficients written with secants:
    K = 2 sec^2(phi) (1 -/+ sqrt|1 - cos^2(phi) sec^2(delta)|) - 1."""
    phi, dlt = math.radians(phi_int_deg), math.radians(phi_bed_deg)
    sec2 = 1.0 / math.cos(phi) ** 2
    root = math.sqrt(abs(1.0 - math.cos(phi) ** 2 / math.cos(dlt) ** 2))
    return 2 * sec2 * (1 - root) - 1, 2 * sec2 * (1 + root) - 1, 2 * sec2 - 1


def _linear_velocity_state(nx, ny, x0, y0, a, b, h0=1.0):
    xc = x0 + np.arange(nx, dtype=np.float64)
    yc = y0 + np.arange(ny, dtype=np.float64)
    X, Y = np.meshgrid(xc, yc)
    h = np.full(nx * ny, h0)
    return [h, (h0 * a * X).reshape(-1), (h0 * b * Y).reshape(-1)], X.reshape(-1), Y.reshape(-1)


def _interior(nx, ny):
    i = np.arange(nx * ny) % nx
    j = np.arange(nx * ny) // nx
    return (i > 0) & (i < nx - 1) & (j > 0) & (j < ny - 1)


@pytest.mark.parametrize("a,b,branch", [(0.25, 0.125, "active"), (-0.25, 0.125, "passive"),
                                         (0.25, -0.25, "neutral"), (-0.5, -0.25, "passive")])
def test_kact_branch_follows_velocity_divergence(oracle_lib, a, b, branch):
    nx = ny = 12
    eps = 0.5
    orc = _flat_oracle(nx, ny, 1.0, -6.0, -6.0, 30.0, 20.0, eps=eps)
    U, X, Y = _linear_velocity_state(nx, ny, -6.0, -6.0, a, b)
    k, d = orc.kact_slopes(U)
    inner = _interior(nx, ny)
    # the slopes of a linear field are exact: div u = a + b
    np.testing.assert_array_equal(d[inner, 1], a)
    np.testing.assert_array_equal(d[inner, 5], b)
    act, pas, neu = _earth_pressure(30.0, 20.0)
    want = {"active": act, "passive": pas, "neutral": neu}[branch]
    rest = inner & (X == 0) & (Y == 0)        # |u|, |v| < tiny: K = 1
    moving = inner & ~rest
    np.testing.assert_allclose(k[moving], eps * want, rtol=1e-13)
    np.testing.assert_array_equal(k[rest], eps * 1.0)
    # boundary cells have zero slopes: neither branch
    ring = ~inner & ~((X == 0) & (Y == 0))
    np.testing.assert_allclose(k[ring], eps * neu, rtol=1e-13)
    assert act < neu < pas


def test_kact_clamp_at_zero_and_dry_cells(oracle_lib):
    # phi_bed > phi_int: |1 - cos^2 phi sec^2 delta| > 0 and the active
    # coefficient is negative -> clamped to 0; the passive one stays positive
    nx = ny = 10
    orc = _flat_oracle(nx, ny, 1.0, -5.0, -5.0, 20.0, 40.0)
    act, pas, _ = _earth_pressure(20.0, 40.0)
    assert act < 0 < pas
    inner = _interior(nx, ny)
    for a, want in ((0.25, 0.0), (-0.25, pas)):
        U, X, Y = _linear_velocity_state(nx, ny, -5.0, -5.0, a, 0.0)
        k, _ = orc.kact_slopes(U)
        sel = inner & (X != 0)
        np.testing.assert_allclose(k[sel], want, rtol=1e-13, atol=0)
    # h <= tiny: K = 1 (times epsilon) whatever the velocities
    U, _, _ = _linear_velocity_state(nx, ny, -5.0, -5.0, 0.25, 0.0, h0=5e-5)
    k, _ = orc.kact_slopes(U)
    assert np.all(k == 1.0)


@pytest.mark.parametrize("gamma,row,want", [
    (1.0, [0.0, 2.0, 6.0], 1.0),          # dp = 2, dm = 1: min |dm|
    (1.0, [0.0, 4.0, 6.0], 1.0),          # dp = 1, dm = 2: min |dp|
    (1.0, [6.0, 4.0, 0.0], -1.0),         # monotone decreasing: sign kept
    (1.0, [0.0, 2.0, 0.0], 0.0),          # extremum: opposite signs
    (1.0, [0.0, 1e-4, 2.0], 2.5e-5),      # |dm| = 5e-5 in the dead zone: 1/2 min(...)
    (1.0, [2.0, 2.0 - 1e-4, 0.0], -2.5e-5),
    (1.0, [0.0, 5e-5, 1e-4], 0.0),        # both in the dead zone
    (2.0, [0.0, 2.0, 6.0], 1.5),          # gamma 2: the central difference is the min
    (2.0, [0.0, 2.0, 2.5], 0.5),          # gamma 2: 2 |dp| = 0.5 < |dc| = 0.625
])
def test_limiter_hand_stencils(oracle_lib, gamma, row, want):
    # three cells in x, spacing 2; the middle cell of the middle row is the
    # only interior cell; its x slope sees exactly this stencil
    nx, ny, dx = 3, 3, 2.0
    orc = _flat_oracle(nx, ny, dx, 0.0, 0.0, 30.0, 20.0, gamma=gamma)
    h = np.full(9, 1.0)
    h[3:6] = row
    U = [h, np.zeros(9), np.zeros(9)]
    _, d = orc.kact_slopes(U)
    assert d[4, 0] == pytest.approx(want, rel=1e-10, abs=1e-300)
    # the y stencil (1, row[1], 1) is an extremum in every case: 0
    assert d[4, 3] == 0.0
    # boundary cells have zero slopes
    assert not d[np.arange(9) != 4].any()


def _sheet_step(theta_deg, phi_bed, u0, kx=0.0, s=60, dt=0.01, h0=1.0):
    th = math.radians(theta_deg)
    nx = ny = 8
    orc = _flat_oracle(nx, ny, 10.0, 0.0, 0.0, 30.0, phi_bed, gx=9.8 * math.sin(th),
                       gz=9.8 * math.cos(th), kx=kx)
    n = nx * ny
    U = [np.full(n, h0), np.full(n, h0 * u0), np.zeros(n)]
    orc.step(U, s, dt)
    return U


@pytest.mark.parametrize("theta,phi_bed,u0", [(30.0, 20.0, 2.0), (20.0, 30.0, 2.0),
                                              (30.0, 20.0, -2.0), (35.0, 22.0, 0.5)])
def test_coulomb_sheet_on_a_plane(oracle_lib, theta, phi_bed, u0):
    # a uniform sheet: every face flux cancels and hu changes by the sources
    # alone: du/dt = g sin(theta) - sgn(u) g cos(theta) tan(phi_bed)
    dt = 0.01
    U = _sheet_step(theta, phi_bed, u0, dt=dt)
    th, pb = math.radians(theta), math.radians(phi_bed)
    acc = 9.8 * math.sin(th) - math.copysign(1.0, u0) * 9.8 * math.cos(th) * math.tan(pb)
    np.testing.assert_allclose(U[1], u0 + dt * acc, rtol=1e-14)
    assert np.all(U[0] == 1.0) and np.all(U[2] == 0.0)
    if phi_bed > theta and u0 > 0:
        assert acc < 0          # friction steeper than the slope decelerates


def test_coulomb_curvature_and_liftoff(oracle_lib):
    dt, u0, theta, pb = 0.01, 2.0, 30.0, 20.0
    th, tb = math.radians(theta), math.tan(math.radians(pb))
    # curvature adds the centripetal load u^2 kappa to the normal force
    kappa = 0.3
    U = _sheet_step(theta, pb, u0, kx=kappa, dt=dt)
    acc = 9.8 * math.sin(th) - (9.8 * math.cos(th) + u0 * u0 * kappa) * tb
    np.testing.assert_allclose(U[1], u0 + dt * acc, rtol=1e-14)
    # a convex bend strong enough to cancel the normal load: no friction
    U = _sheet_step(theta, pb, u0, kx=-9.8, dt=dt)
    np.testing.assert_allclose(U[1], u0 + dt * 9.8 * math.sin(th), rtol=1e-14)


def test_coulomb_dead_zone_and_rest(oracle_lib):
    dt, theta = 0.01, 30.0
    g_sin = 9.8 * math.sin(math.radians(theta))
    # steps 0..49 use frict_tiny = 0.1: |u| = 0.05 slides frictionless
    U = _sheet_step(theta, 20.0, 0.05, s=10, dt=dt)
    np.testing.assert_allclose(U[1], 0.05 + dt * g_sin, rtol=1e-14)
    # after the switch the same speed feels the bed
    U = _sheet_step(theta, 20.0, 0.05, s=60, dt=dt)
    acc = g_sin - 9.8 * math.cos(math.radians(theta)) * math.tan(math.radians(20.0))
    np.testing.assert_allclose(U[1], 0.05 + dt * acc, rtol=1e-14)
    # at rest tiny_sgn(0) = 0: the scheme has no static friction, a resting
    # sheet starts to slide even with phi_bed above the slope angle
    U = _sheet_step(theta, 40.0, 0.0, s=60, dt=dt)
    np.testing.assert_allclose(U[1], dt * g_sin, rtol=1e-14)


def test_curvature_of_a_quadratic_dem():
    # z = a X^2 + b Y^2 on a raster of spacing 2 with dyadic values: every
    # Horn slope and curvature difference is exact, so the interior curvature
    # is (2a, 2b) exactly.  Edge slopes copy their inner neighbour, so a
    # difference that reaches the first or last slope column / row sees half
    # the change: the direct 3x3 branch (last three columns / rows) gives a
    # (or b) next to an edge, and the interpolated branch blends a and 2a with
    # its weight x1 = (x - xmin - res) / res in (-1, 1) in the first band
    # (the reference moves row / column 0 to 1 first)
    a, b = 1.0 / 64, 1.0 / 32
    res, nr, nc = 2.0, 33, 39
    xmin, ymax = 0.0, (nr - 1) * res
    X, Y = np.meshgrid(xmin + np.arange(nc) * res, ymax - np.arange(nr) * res)
    elev = np.ascontiguousarray((a * X * X + b * Y * Y).astype(np.float32).reshape(-1))
    assert np.array_equal(elev.astype(np.float64), (a * X * X + b * Y * Y).reshape(-1))
    nx, ny, dx = 25, 21, 3.1
    x0 = y0 = 0.5
    _, _, cv = problems.static_coefs((nr, nc, xmin, ymax, res, elev), nx, ny, dx, dx, x0, y0)
    cx, cy = cv[0].reshape(ny, nx), cv[1].reshape(ny, nx)
    xs = x0 + np.arange(nx) * dx
    ys = y0 + np.arange(ny) * dx
    seen = set()
    for j in range(ny):
        row = int((ymax - ys[j]) / res)
        row = min(max(row, 1), nr - 2)
        for i in range(nx):
            col = int((xs[i] - xmin) / res)
            col = min(max(col, 1), nc - 2)
            direct = col >= nc - 3 or row >= nr - 3
            if direct:
                wx = a if col in (1, nc - 2) else 2 * a
                wy = b if row in (1, nr - 2) else 2 * b
            else:
                wx = a * (1.0 + (xs[i] - xmin - res) / res) if col == 1 else 2 * a
                wy = b * (1.0 + (ymax - res - ys[j]) / res) if row == 1 else 2 * b
            seen.add((direct, wx == 2 * a, wy == 2 * b))
            assert cx[j, i] == pytest.approx(wx, rel=1e-12, abs=1e-15), (i, j)
            assert cy[j, i] == pytest.approx(wy, rel=1e-12, abs=1e-15), (i, j)
    # interior, halved and interpolated edge values all occur
    assert len(seen) >= 5


def _exact_bucket(k0, k1, R, nb):
    v = Fraction(k0 * R + k1) * Fraction(8 * nb, R * R + R) + Fraction(1, 2)
    return (v.numerator // v.denominator) % nb, v - (v.numerator // v.denominator)


def test_hash_bucket_at_boundaries():
    # hash = (int)((k0 R + k1) * 8 NB / (R R + R) + 1/2) % NB with R = 2^32 + 1
    assert mesh.KEY_RANGE == 2 ** 32 + 1
    R = 2 ** 32 + 1
    nb = 100000
    # hand-worked: key 0 -> 0; the top of wrap 0 at k0 = 2^29 rounds into the
    # first bucket of the next wrap, i.e. 0 again; k0 = 2^29 - 2^13 lands
    # 8 NB 2^13 / (2^32 + 2) ~ 1.526 buckets below the wrap: NB - 2
    keys = [(0, 0), (1 << 29, 0), ((1 << 29) - (1 << 13), 0), (0, 2 ** 32 - 1)]
    want = [0, 0, nb - 2, 0]
    got = mesh.hash_bucket(np.array(keys, dtype=np.uint32), nb).tolist()
    assert got == want
    # keys just below / above the rounding boundary of a bucket in each of the
    # eight wraps: distance to the half-integer between 1e-8 and 2e-5, so the
    # double evaluation agrees with exact arithmetic; an R R denominator (no
    # + R) or a missing + 1/2 would move these keys to the next bucket
    rng = np.random.default_rng(11)
    keys, want = [], []
    scale = Fraction(R * R + R, 8 * nb)            # key units per bucket
    for wrap in range(8):
        for _ in range(6):
            bkt = int(rng.integers(1, nb - 1))
            target = (wrap * nb + bkt - Fraction(1, 2)) * scale   # value = ...+0.5 boundary
            for side in (-1, 1):
                off = Fraction(int(rng.integers(5, 1000)) * int(side), 100000000) * scale
                K = int(target + off)
                k0, k1 = divmod(K, R)
                if k1 > 2 ** 32 - 1:
                    continue
                exp, frac = _exact_bucket(k0, k1, R, nb)
                dist = min(frac, 1 - frac)
                if not (Fraction(1, 10 ** 8) < dist < Fraction(2, 10 ** 5)):
                    continue
                keys.append((k0, k1))
                want.append(exp)
                # the two sides of a boundary differ by one bucket
                assert exp == (bkt - (1 if side < 0 else 0)) % nb
    assert len(keys) >= 80
    got = mesh.hash_bucket(np.array(keys, dtype=np.uint32), nb).tolist()
    assert got == want
