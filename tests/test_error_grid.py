"""Error grid (CPU): the oracle's pieces pinned by closed forms and a LAPACK
solve, and the product's host stencils (tg_host_error_stencil) against the
oracle's geometric restatement on uniform and 2:1 adaptive meshes.

Pins: the 4x4 solve against numpy.linalg.solve (LAPACK dgesv) within the
conditioning of the system; the four interpolation rules reproduce the
fields they are exact for (bilinear, affine) and linear_interp's missing
constant term is kept; interior stencils of a uniform mesh are the father,
its side neighbours and the diagonal one, edge stencils fall to two points
and corner stencils to one; the error-grid update equals the forward
step's on the same grid up to the slopes' reciprocal rounding; a lake at
rest has zero residual; a constant adjoint has zero error where the rule
reproduces constants and |res . lambda| where linear_interp drops them.
"""
import numpy as np
import pytest

from discrete_adjoint_amd import errorgrid, mesh
from helpers import analytic_leaf_problem, disc, paraboloid
import oracle
from oracle import error_oracle as eo


def test_dgesv4_matches_lapack_within_conditioning():
    rng = np.random.default_rng(3)
    for _ in range(200):
        A = rng.normal(size=(4, 4))
        b = rng.normal(size=4)
        x, sing = eo.dgesv4(A, b)
        assert not sing
        ref = np.linalg.solve(A, b)
        tol = 1e-15 * np.linalg.cond(A) * 10
        assert np.allclose(x, ref, rtol=tol, atol=tol * np.abs(ref).max())
    # bilinear fits at realistic coordinates (centres of 5 m cells near 2 km)
    for _ in range(50):
        x0, y0 = rng.uniform(1000, 3000, 2)
        d = 5.0 * rng.choice([1, 2])
        pts = [(x0, y0), (x0 + d, y0), (x0, y0 + d), (x0 + d, y0 + d)]
        A = np.array([[1, px, py, px * py] for px, py in pts])
        f = rng.normal(size=4)
        x, _ = eo.dgesv4(A, f)
        ref = np.linalg.solve(A, f)
        tol = 1e-15 * np.linalg.cond(A) * 10
        assert np.allclose(x, ref, rtol=tol, atol=tol * np.abs(ref).max())
    # an exactly zero pivot: dgesv leaves the right side as it is
    A = [[1, 0, 0, 0], [1, 1, 0, 0], [1, 2, 0, 0], [1, 0, 1, 0]]
    b, sing = eo.dgesv4(A, [1.0, 2.0, 3.0, 4.0])
    assert sing and b == [1.0, 2.0, 3.0, 4.0]


def test_interpolation_rules_reproduce_their_fields():
    rng = np.random.default_rng(5)
    a0, a1, a2, a3 = rng.normal(size=4)

    def bil(x, y):
        return a0 + a1 * x + a2 * y + a3 * x * y
    u = [(0.0, 0.0, 0), (2.0, 0.0, 1), (0.0, 2.0, 2), (2.0, 2.0, 3)]
    for x, y in rng.uniform(0, 2, (20, 2)):
        got = eo.interpolate(eo.BILINEAR, u, [bil(p[0], p[1]) for p in u], x, y)
        assert got == pytest.approx(bil(x, y), rel=1e-12, abs=1e-12)

    def aff(x, y):
        return a0 + a1 * x + a2 * y
    u3 = [(0.0, 0.0, 0), (2.0, 0.0, 1), (0.0, 2.0, 2)]
    for x, y in rng.uniform(0, 2, (20, 2)):
        got = eo.interpolate(eo.BARYCENTRIC, u3, [aff(p[0], p[1]) for p in u3], x, y)
        assert got == pytest.approx(aff(x, y), rel=1e-12, abs=1e-12)
    # linear_interp as written: (f2 - f1) / (x2 - x1) * (x - x1), no f1 term
    u2 = [(1.0, 5.0, 0), (3.0, 5.0, 1)]
    assert eo.interpolate(eo.LINEAR_X, u2, [10.0, 14.0], 1.5, 5.0) == 1.0
    u2y = [(1.0, 5.0, 0), (1.0, 9.0, 1)]
    assert eo.interpolate(eo.LINEAR_Y, u2y, [10.0, 2.0], 1.0, 6.0) == -2.0
    assert eo.interpolate(eo.COPY, u2[:1], [7.5], 0.3, 0.4) == 7.5
    # all four values zero: zero without a solve
    assert eo.interpolate(eo.BILINEAR, u, [0.0] * 4, 0.5, 0.5) == 0.0


def _grid_of(nx0, ny0, dx0, cells):
    geom = (nx0, ny0, 0.0, 0.0, dx0, dx0)
    return geom, eo.Grid(geom, cells)


def _product_grid(geom, cells):
    def statics(xc):
        n = len(xc)
        return np.tile([[0.0], [0.0], [1.0]], (1, n)), np.zeros((2, n))
    return errorgrid.build(geom, cells, statics)


def _table_cells(eg, geom):
    lev, ix, iy = errorgrid.cells_of_table(eg.table, geom[2], geom[3])
    return list(zip(lev.tolist(), ix.tolist(), iy.tolist()))


def test_uniform_stencils_closed_form():
    nx0, ny0, d = 6, 5, 10.0
    cells = errorgrid.uniform_cells(nx0, ny0)
    geom, g = _grid_of(nx0, ny0, d, cells)
    eg = _product_grid(geom, cells)
    kinds = {}
    for q, c in enumerate(_table_cells(eg, geom)):
        L, i, j = c
        I, J = i >> 1, j >> 1
        a, b = i & 1, j & 1
        # the father and its neighbours towards the son's corner
        sx, sy = (1 if a else -1), (1 if b else -1)
        want = {(I, J), (I + sx, J), (I, J + sy), (I + sx, J + sy)}
        want = {p for p in want if 0 <= p[0] < nx0 and 0 <= p[1] < ny0}
        got = {(int(p) % nx0, int(p) // nx0) for p in eg.pts[q][eg.pts[q] >= 0]}
        assert got == want, (c, got, want)
        n_pts = len(want)
        expect = {1: eo.COPY, 2: None, 4: eo.BILINEAR}[n_pts] if n_pts != 3 else None
        if n_pts == 2:
            expect = eo.LINEAR_X if len({p[1] for p in want}) == 1 else eo.LINEAR_Y
        assert eg.kind[q] == expect, c
        kinds[int(eg.kind[q])] = kinds.get(int(eg.kind[q]), 0) + 1
        # points are the fathers' centres
        for k in range(n_pts):
            f = eg.pts[q][k]
            assert np.allclose(eg.pxy[q][k], [(f % nx0 + .5) * d, (f // nx0 + .5) * d], atol=1e-12)
    assert kinds[eo.COPY] == 4 and kinds[eo.BILINEAR] == (2 * nx0 - 2) * (2 * ny0 - 2)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_product_stencils_equal_the_oracle_on_adaptive_meshes(seed):
    nx0, ny0, d = 7, 6, 8.0
    cells = mesh.random_balanced_tree(nx0, ny0, 3, 0.35, seed)
    geom, g = _grid_of(nx0, ny0, d, cells)
    eg = _product_grid(geom, cells)
    tc = _table_cells(eg, geom)
    seen = set()
    for q, c in enumerate(tc):
        kind, u = g.stencil(c)
        assert eg.kind[q] == kind, c
        assert eg.father[q] == g.el[c][0] and eg.which_son[q] == g.el[c][1]
        for k in range(4):
            if k < len(u):
                assert eg.pts[q][k] == u[k][2], (c, k)
                assert eg.pxy[q][k][0] == u[k][0] and eg.pxy[q][k][1] == u[k][1]
            else:
                assert eg.pts[q][k] == -1
        seen.add(kind)
    # every rule occurs: corners copy, edges interpolate linearly, a
    # coarser neighbour merges two points into one (barycentric)
    assert seen == {eo.COPY, eo.LINEAR_X, eo.LINEAR_Y, eo.BARYCENTRIC, eo.BILINEAR}


def _leaf_case(seed=4, nsteps=3):
    """A small adaptive flow, its step-s states from the leaf oracle, an
    adjoint-like field, the product's error grid and the oracle's."""
    nx0, ny0, d = 12, 10, 20.0
    pb = analytic_leaf_problem(nx0, ny0, d, 2, disc(0.4 * nx0, 0.5 * ny0, 2.5), slope_x_deg=22.0,
                               slope_y_deg=4.0, h=paraboloid(0.4 * nx0 * d, 0.5 * ny0 * d, 45.0,
                                                             35.0, 6.0),
                               phi_bed=16.0, curv=(1e-3, -5e-4))
    orc = oracle.LeafOracle.from_problem(pb)
    U = [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]
    for s in range(nsteps):
        prev = [x.copy() for x in U]
        dt, _ = orc.step(U, s)
    rng = np.random.default_rng(seed)
    X, Y = pb.table.xc[:, 0], pb.table.xc[:, 1]
    lam = np.stack([1.0 + 0.001 * X + 0.002 * Y, 0.01 * np.sin(X / 50.0),
                    0.01 * np.cos(Y / 40.0)]) + 1e-3 * rng.normal(size=(3, pb.n))
    cells = errorgrid.cells_of_table(pb.table, pb.xmin, pb.ymin)
    geom = (pb.nx0, pb.ny0, pb.xmin, pb.ymin, pb.dx0, pb.dy0)

    def statics(xc):
        n = len(xc)
        return np.repeat(pb.grav[:, :1], n, axis=1), np.repeat(pb.curv[:, :1], n, axis=1)
    eg = errorgrid.build(geom, cells, statics)
    return pb, np.array(prev), np.array(U), lam, nsteps - 1, dt, eg, geom, cells


def error_oracle_of(pb, eg, geom, cells, prev, state, lam, s, dt):
    """The oracle's reconstruction, residual, error and correction on the
    product's error grid order."""
    g = eo.Grid(geom, cells)
    tc = _table_cells(eg, geom)
    fwd = np.concatenate([prev, state, lam])
    rec, _ = eo.reconstruct(g, tc, fwd)
    t = eg.table
    lo = oracle.LeafOracle(t.nbr, t.nkind, t.gen, t.dx, t.xc, eg.grav, eg.dgrav, eg.curv,
                           pb.phi_int_deg, pb.phi_bed_deg, pb.epsilon, pb.gravity_scale,
                           mat=eg.mat, materials=getattr(pb, "materials", None))
    res, err, corr = eo.indicator(lo, rec, lam[:, eg.father], s, dt)
    return rec, res, err, corr, lo


def test_error_update_equals_forward_step_up_to_slope_rounding(oracle_lib):
    pb, prev, state, lam, s, dt, eg, geom, cells = _leaf_case()
    rec, res, err, corr, lo = error_oracle_of(pb, eg, geom, cells, prev, state, lam, s, dt)
    U = [rec[0].copy(), rec[1].copy(), rec[2].copy()]
    lo.step(U, s, dt)
    pred = lo.error_update(rec[0:3], s, dt)
    for a, b in zip(pred, U):
        assert np.allclose(a, b, rtol=1e-11, atol=1e-11 * np.abs(b).max())
    assert np.abs(res).max() > 0 and np.isfinite(err).all()


def test_lake_at_rest_has_zero_residual(oracle_lib):
    nx0, ny0, d = 8, 6, 10.0
    pb = analytic_leaf_problem(nx0, ny0, d, 2, disc(4, 3, 1.5), h=lambda X, Y: 2.0 + 0 * X)
    cells = errorgrid.cells_of_table(pb.table, 0.0, 0.0)
    geom = (nx0, ny0, 0.0, 0.0, d, d)

    def statics(xc):
        n = len(xc)
        return np.repeat(pb.grav[:, :1], n, axis=1), np.zeros((2, n))
    eg = errorgrid.build(geom, cells, statics)
    U = np.array([pb.h, pb.hu, pb.hv])
    lam = np.ones((3, pb.n))
    rec, res, err, corr, _ = error_oracle_of(pb, eg, geom, cells, U, U, lam, 0, 0.1)
    # where the rule reproduces constants the reconstructions are exact and
    # the residual is zero; linear_interp's rows (edges) drop the constant
    lin = (eg.kind == eo.LINEAR_X) | (eg.kind == eo.LINEAR_Y)
    assert lin.any() and (~lin).any()
    assert np.allclose(rec[:, ~lin], np.repeat(U[:, :1], 3, axis=0)[:, :1] * 0 +
                       np.concatenate([U, U, lam])[:, eg.father][:, ~lin], rtol=1e-12)
    assert np.all(rec[:, lin] == 0.0)
    # away from them (two rings: fluxes and the neighbours' kact) the lake
    # stays at rest
    near = lin.copy()
    nb = eg.table.nbr
    for _ in range(2):
        grow = near.copy()
        for k in range(8):
            ok = nb[:, k] >= 0
            grow[ok] |= near[nb[ok, k]]
        near = grow
    interior = ~near & (eg.kind != eo.COPY)
    assert interior.sum() > eg.n // 3
    assert np.abs(res[:, interior]).max() < 1e-12
    assert np.abs(err[interior]).max() < 1e-12


def test_constant_adjoint_error_marks_linear_rows(oracle_lib):
    pb, prev, state, _, s, dt, eg, geom, cells = _leaf_case()
    lam = np.full((3, pb.n), 2.0)
    rec, res, err, corr, _ = error_oracle_of(pb, eg, geom, cells, prev, state, lam, s, dt)
    lin = (eg.kind == eo.LINEAR_X) | (eg.kind == eo.LINEAR_Y)
    exact = (eg.kind == eo.COPY) | (eg.kind == eo.BILINEAR) | (eg.kind == eo.BARYCENTRIC)
    assert np.allclose(err[exact], 0.0, atol=1e-10 * np.abs(res).max())
    assert np.allclose(err[lin], 2.0 * np.abs(res[:, lin]).sum(axis=0), rtol=1e-14)
    assert np.allclose(corr, 2.0 * res.sum(axis=0), rtol=1e-12, atol=1e-300)
