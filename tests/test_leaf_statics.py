"""CPU: d gravity_z on adaptive meshes (calc_d_gravity, element2.C:3574-3653)
through tg_host_leaf_d_gravity.  Pinned three ways: on a single-level tree it
equals the uniform-block coefficients (tg_host_static_coefs, itself checked
in closed form); a gz linear along one direction gives that slope at every
leaf, hanging sides included (so the centre distances are the right ones);
on random 2:1-balanced trees the neighbour each side uses is the one a
geometric search finds next to corner s (neighbor[s], not neighbor[s + 4])."""
import numpy as np
import pytest

from discrete_adjoint_amd import _lib, mesh
from discrete_adjoint_amd.problems import static_coefs


def _tree(seed, nx0=6, ny0=5, max_level=3, frac=0.35, dx0=8.0, dy0=6.0, x0=100.0, y0=-50.0):
    lev, ix, iy = mesh.random_balanced_tree(nx0, ny0, max_level, frac, seed)
    t = mesh.flatten(nx0, ny0, x0, y0, dx0, dy0, lev, ix, iy)
    return t, (lev, ix, iy), (nx0, ny0, dx0, dy0, x0, y0)


def test_single_level_equals_uniform_block():
    nx, ny, s = 12, 9, 10.0
    lev = np.zeros(nx * ny, np.int32)
    iy, ix = np.divmod(np.arange(nx * ny, dtype=np.int32), nx)
    t = mesh.flatten(nx, ny, 0.0, 0.0, s, s, lev, ix, iy)
    rng = np.random.default_rng(4)
    elev = (rng.standard_normal((40, 40)) * 3.0).astype(np.float32).reshape(-1)
    dem = (40, 40, -60.0, 200.0, 6.5, elev)
    g, dg, _ = static_coefs(dem, nx, ny, s, s, 0.5 * s, 0.5 * s)
    cell = iy[t.order].astype(np.int64) * nx + ix[t.order]
    got = mesh.leaf_d_gravity(t, g[2, cell])
    # centre differences are exact multiples of s here: the same values
    assert np.allclose(got, dg[:, cell], rtol=1e-13, atol=0.0)


@pytest.mark.parametrize("seed", [1, 7])
def test_linear_gz_gives_its_slope_on_every_leaf(seed):
    t, _, _ = _tree(seed)
    x, y = t.xc[:, 0], t.xc[:, 1]
    gx = 9.0 + 2.5e-3 * x                     # varies along x only
    d = mesh.leaf_d_gravity(t, gx)
    hasx = (t.nbr[:, 1] >= 0) | (t.nbr[:, 3] >= 0)
    assert hasx.all()
    assert np.allclose(d[0], 2.5e-3, rtol=1e-9, atol=0.0)
    # along y: 0 where the y neighbours sit straight above / below; a finer
    # neighbour across a hanging side is off-centre in x, and the reference's
    # difference ignores that offset
    _off_centre(t, d[1], axis=1)
    gy = 9.0 - 4e-3 * y
    d = mesh.leaf_d_gravity(t, gy)
    assert np.allclose(d[1], -4e-3, rtol=1e-9, atol=0.0)
    _off_centre(t, d[0], axis=0)


def _off_centre(t, d, axis):
    other = 1 - axis
    slots = (2, 0) if axis == 1 else (1, 3)
    straight = np.ones(len(d), bool)
    for s in slots:
        nb = t.nbr[:, s]
        has = nb >= 0
        straight[has] &= t.xc[nb[has], other] == t.xc[has, other]
    assert not np.any(d[straight])
    assert np.any(d[~straight] != 0)


def _slot_neighbour(leaf_at, L, i, j, side, nx0, ny0):
    """The leaf next to corner `side` of cell (L, i, j) across side `side`
    (0 S, 1 E, 2 N, 3 W; corners 0 SW, 1 SE, 2 NE, 3 NW), found by point
    location on the finest grid; -1 outside the domain."""
    F = 8  # integer coordinates with 2^8 points per base cell
    u = 1 << (F - L)
    X0, Y0 = i * u, j * u
    px, py = {0: (X0, Y0 - 1), 1: (X0 + u, Y0), 2: (X0 + u - 1, Y0 + u),
              3: (X0 - 1, Y0 + u - 1)}[side]
    if px < 0 or py < 0 or px >= nx0 << F or py >= ny0 << F:
        return -1
    for Lq in range(F):
        k = (Lq, px >> (F - Lq), py >> (F - Lq))
        if k in leaf_at:
            return leaf_at[k]
    raise AssertionError("no leaf holds the point")


@pytest.mark.parametrize("seed", [2, 3, 11])
def test_neighbours_are_the_slot_s_leaves(seed):
    t, (lev, ix, iy), (nx0, ny0, *_r) = _tree(seed)
    n = len(lev)
    pos = np.empty(n, np.int64)
    pos[t.order] = np.arange(n)
    leaf_at = {(int(lev[e]), int(ix[e]), int(iy[e])): int(pos[e]) for e in range(n)}
    rng = np.random.default_rng(seed)
    gz = 9.8 - rng.uniform(0.0, 0.3, n)
    got = mesh.leaf_d_gravity(t, gz)
    for q in range(n):
        e = t.order[q]
        nb = {s: _slot_neighbour(leaf_at, int(lev[e]), int(ix[e]), int(iy[e]), s, nx0, ny0)
              for s in range(4)}
        for s in range(4):
            assert nb[s] == t.nbr[q, s], (q, s)
        for d, (p, m) in enumerate(((1, 3), (2, 0))):
            ep, em, c = nb[p], nb[m], t.xc[q, d]
            if ep >= 0 and em >= 0:
                dxp, dxm = t.xc[ep, d] - c, c - t.xc[em, d]
                dp, dm = (gz[ep] - gz[q]) / dxp, (gz[q] - gz[em]) / dxm
                ref = (dp * dxm + dm * dxp) / (dxm + dxp)
            elif em >= 0:
                ref = (gz[q] - gz[em]) / (c - t.xc[em, d])
            elif ep >= 0:
                ref = (gz[ep] - gz[q]) / (t.xc[ep, d] - c)
            else:
                ref = 0.0
            assert got[d, q] == ref, (q, d)


def test_bad_tables_are_rejected():
    t, _, _ = _tree(1)
    nbr = np.array(t.nbr, dtype=np.int32)
    nbr[3, 2] = len(nbr)
    out = np.zeros(2 * len(nbr))
    gz = np.ones(len(nbr))
    xc = np.ascontiguousarray(t.xc).reshape(-1)
    rc = _lib.load().tg_host_leaf_d_gravity(len(nbr), _lib.iptr(nbr.reshape(-1)), _lib.dptr(xc),
                                            _lib.dptr(gz), _lib.dptr(out))
    assert rc == -1
