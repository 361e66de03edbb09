"""Pins the adaptive-mesh CPU restatement (oracle/leaf_oracle.c) without
retyping its formulas:

* on a mesh of equal leaves it reduces bit for bit to the uniform-block
  restatement (forward states, dt, Jacobian blocks, adjoint, bed-friction
  gradient), which its own tests pin;
* hanging sides: mass conservation to rounding (the half-sum big-face flux
  against the two small faces), lake at rest and mirror symmetry to the last
  bit, limited slopes exact for fields linear in x (the two-neighbour
  average and centre-distance differences of element2.C:886-897);
* adjoint: Jacobian blocks, including slots 5..8 of hanging sides, against
  central finite differences of the forward step; the dot-product identity
  of the which_neighbor gather; lambda^0 and dJ/dphi_bed against finite
  differences of whole runs at frozen dt.
"""
import math

import numpy as np
import pytest

from helpers import (analytic_leaf_problem, analytic_problem, disc, paraboloid,
                     small_flow_problem, small_leaf_flow_problem)


def _U(pb):
    return [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]


def _leaf(oracle_lib, pb):
    return oracle_lib.LeafOracle.from_problem(pb)


def _uniform_pair(nsteps=30):
    """small_flow_problem and the same 24 x 20 grid as a leaf table (one
    level), with the leaf -> uniform cell map."""
    upb = small_flow_problem(nsteps=nsteps)
    L = upb.nx * upb.dx
    lpb = analytic_leaf_problem(upb.nx, upb.ny, upb.dx, 0, lambda L_, I, J: I < 0,
                                slope_x_deg=25.0, slope_y_deg=3.0,
                                h=paraboloid(0.35 * L, 0.5 * upb.ny * 10.0, 60.0, 45.0, 8.0),
                                phi_int=32.0, phi_bed=15.0, nsteps=nsteps, curv=(1e-3, -5e-4))
    cell = lpb.table.order  # base leaves are listed row by row
    assert np.array_equal(upb.h[cell], lpb.h)
    return upb, lpb, cell


def test_equal_leaves_reduce_to_the_uniform_restatement(oracle_lib):
    upb, lpb, cell = _uniform_pair(40)
    uo, lo = oracle_lib.Oracle.from_problem(upb), _leaf(oracle_lib, lpb)
    Uu, Ul = _U(upb), _U(lpb)
    traj = []
    for s in range(40):
        traj.append(([a.copy() for a in Uu], [a.copy() for a in Ul]))
        du, mu = uo.step(Uu, s)
        dl, ml = lo.step(Ul, s)
        assert du == dl
        for k in range(3):
            assert np.array_equal(Uu[k][cell], Ul[k])
    # reverse step at a moving state: blocks, adjoint, gradient
    (Su, Sl), dt = traj[30], 0.05
    Ju, Jl = uo.jacobian(Su, 30, dt), lo.jacobian(Sl, 30, dt)
    # uniform slots 1..4 are E, N, W, S; leaf slots 1..4 are S, E, N, W
    assert np.array_equal(Ju[cell][:, [0, 4, 1, 2, 3]], Jl[:, :5])
    assert not Jl[:, 5:].any()
    rng = np.random.default_rng(4)
    lam = rng.standard_normal((3, upb.n))
    lu, gu = uo.adjoint_step(Su, 30, dt, lam, want_dphi=True)
    ll, gl = lo.adjoint_step(Sl, 30, dt, lam[:, cell], want_dphi=True)
    assert np.array_equal(lu[:, cell], ll)
    assert np.array_equal(uo.dstep_dphi(Su, 30, dt)[:, cell], lo.dstep_dphi(Sl, 30, dt))
    assert gl == pytest.approx(gu, rel=1e-12)


def test_mass_is_conserved_across_hanging_sides(oracle_lib):
    # boundary leaves pass on whatever enters them (outflow copy), so mass
    # is conserved while the front stays off the outer two rings
    pb = small_leaf_flow_problem(nsteps=45)
    t = pb.table
    assert (t.nkind[:, 4:] >= 0).sum() > 50      # hanging sides present
    area = t.dx[:, 0] * t.dx[:, 1]
    m0 = float(np.sum(area * pb.h))
    orc = _leaf(oracle_lib, pb)
    U = _U(pb)
    for s in range(45):
        _, mass = orc.step(U, s)
        assert mass == pytest.approx(m0, rel=1e-13)
        assert (U[0] >= 0).all()
    x, y = t.xc[:, 0], t.xc[:, 1]
    outer = (x < 40) | (x > 280) | (y < 40) | (y > 200)
    assert U[0][outer].max() < 1e-3
    # the front has crossed hanging sides: wet leaves of every level
    assert set(t.gen[U[0] > 1e-3]) == {0, 1, 2}
    assert float(np.sum(area * U[0])) == pytest.approx(m0, rel=1e-13)


def test_lake_at_rest_on_an_adaptive_mesh_is_exact(oracle_lib):
    pb = analytic_leaf_problem(12, 10, 20.0, 2, disc(5.0, 5.0, 2.5),
                               h=lambda X, Y: np.full(X.shape, 2.5))
    orc = _leaf(oracle_lib, pb)
    U = _U(pb)
    for s in range(5):
        orc.step(U, s)
    assert np.array_equal(U[0], pb.h) and not U[1].any() and not U[2].any()


def test_mirror_symmetry_across_hanging_sides(oracle_lib):
    # mesh, ground and pile symmetric about y = Ly/2: h and hu stay
    # symmetric, hv antisymmetric, bit for bit
    nx0, ny0, dx0 = 14, 10, 20.0
    pb = analytic_leaf_problem(nx0, ny0, dx0, 2, disc(5.0, 5.0, 2.2), slope_x_deg=20.0,
                               h=paraboloid(100.0, 100.0, 50.0, 35.0, 6.0), phi_int=32.0,
                               phi_bed=14.0)
    t = pb.table
    key = {(round(x, 9), round(y, 9)): i for i, (x, y) in enumerate(t.xc)}
    mir = np.array([key[(round(x, 9), round(ny0 * dx0 - y, 9))] for x, y in t.xc])
    orc = _leaf(oracle_lib, pb)
    U = _U(pb)
    for s in range(25):
        orc.step(U, s)
    h, hu, hv = U
    assert np.array_equal(h, h[mir]) and np.array_equal(hu, hu[mir])
    assert np.array_equal(hv, -hv[mir])
    assert np.abs(hv).max() > 0


def test_slopes_exact_for_fields_linear_in_x(oracle_lib):
    """u = a x + b y (dyadic a, b, centres): leaves whose x neighbours are
    equal or finer get d/dx = a exactly -- the two finer neighbours of a side
    are averaged (their y offsets cancel), every difference is over the
    centre distance to the first neighbour.  y slopes likewise get b."""
    a, b = 0.25, 0.125
    pb = analytic_leaf_problem(12, 10, 16.0, 2, disc(6.0, 5.0, 2.5),
                               h=lambda X, Y: 40.0 + a * X + b * Y,
                               hu=lambda X, Y: 10.0 + 2 * a * X + b * Y,
                               hv=lambda X, Y: -5.0 + a * X - 2 * b * Y)
    t = pb.table
    orc = _leaf(oracle_lib, pb)
    d = orc.slopes(_U(pb))
    interior = (t.nkind[:, :4] != -1).all(axis=1)
    coarser = t.ngen[:, :4] < t.gen[:, None]
    xok = interior & ~coarser[:, 1] & ~coarser[:, 3]
    yok = interior & ~coarser[:, 0] & ~coarser[:, 2]
    finer_x = xok & ((t.nkind[:, 5] >= 0) | (t.nkind[:, 7] >= 0))
    assert finer_x.sum() >= 8       # averaged sides are covered
    assert np.array_equal(d[xok, 0], np.full(xok.sum(), a))
    assert np.array_equal(d[xok, 1], np.full(xok.sum(), 2 * a))
    assert np.array_equal(d[xok, 2], np.full(xok.sum(), a))
    assert np.array_equal(d[yok, 3], np.full(yok.sum(), b))
    assert np.array_equal(d[yok, 4], np.full(yok.sum(), b))
    assert np.array_equal(d[yok, 5], np.full(yok.sum(), -2 * b))
    assert not d[~interior].any()


def _advance(orc, U, nsteps):
    dts = []
    for s in range(nsteps):
        dt, _ = orc.step(U, s)
        dts.append(dt)
    return dts


def _fd_step(orc, U, s, dt, c, comp, eps):
    outs = []
    for sgn in (+1, -1):
        V = [x.copy() for x in U]
        V[comp][c] += sgn * eps
        t = orc.time
        orc.step(V, s, dt=dt)
        orc.time = t
        outs.append(np.array(V))
    return (outs[0] - outs[1]) / (2 * eps)


def _which(t, f, e):
    for i in range(8):
        if t.nbr[f, i] == e and t.nkind[f, i] >= 0:
            return i
    raise AssertionError


def test_jacobian_blocks_match_finite_differences_on_hanging_sides(oracle_lib):
    pb = small_leaf_flow_problem(nsteps=20)
    t = pb.table
    orc = _leaf(oracle_lib, pb)
    U = _U(pb)
    dts = _advance(orc, U, 20)
    s, dt = 20, dts[-1]
    J = orc.jacobian(U, s, dt)
    wet = np.flatnonzero(U[0] > 0.05)
    # leaves next to a hanging side first, then the rest
    hang = [c for c in wet if (t.ngen[c, :4] != t.gen[c]).any() and (t.nkind[c, :4] != -1).all()]
    rng = np.random.default_rng(0)
    pick = list(hang) + list(rng.choice(wet, size=20, replace=False))
    checked = agree = hang_slots = 0
    for c in pick:
        for comp in range(3):
            e1 = 1e-6 * max(abs(U[comp][c]), 1e-2)
            fd1 = _fd_step(orc, U, s, dt, c, comp, e1)
            fd2 = _fd_step(orc, U, s, dt, c, comp, 2 * e1)
            # leaves whose update depends on U_c: c (slot 0) and every leaf
            # that lists c, in the slot where it lists it
            deps = [(c, 0)]
            for k in range(8):
                f = t.nbr[c, k]
                if f >= 0 and t.nkind[c, k] >= 0:
                    deps.append((f, _which(t, f, c) + 1))
            for j, slot in set(deps):
                an = -J[j, slot][:, comp]
                f1, f2 = fd1[:, j], fd2[:, j]
                if not np.allclose(f1, f2, rtol=1e-4, atol=1e-9):
                    continue
                checked += 1
                if np.allclose(an, f1, rtol=1e-5, atol=1e-7 * max(1.0, np.abs(f1).max())):
                    agree += 1
                    hang_slots += slot > 4
    assert checked > 300
    assert hang_slots > 10
    assert agree >= 0.97 * checked, (agree, checked)


def _tangent(J, dU, t):
    """delta U^{s+1}_j = -sum_slots J_j[slot] delta U^s_{nbr(j, slot)}."""
    n = len(t.gen)
    out = np.zeros((3, n))
    out -= np.einsum("kab,bk->ak", J[:, 0], dU)
    for k in range(8):
        m = (t.nkind[:, k] >= 0)
        src = dU[:, t.nbr[m, k]]
        out[:, m] -= np.einsum("kab,bk->ak", J[m, k + 1], src)
    return out


def test_adjoint_dot_product_identity_with_hanging_slots(oracle_lib):
    pb = small_leaf_flow_problem(nsteps=15)
    orc = _leaf(oracle_lib, pb)
    U = _U(pb)
    dts = _advance(orc, U, 15)
    dt = dts[-1]
    J = orc.jacobian(U, 15, dt)
    assert np.abs(J[:, 5:]).max() > 0
    rng = np.random.default_rng(1)
    dU = rng.standard_normal((3, pb.n))
    lam = rng.standard_normal((3, pb.n))
    lhs = np.sum(lam * _tangent(J, dU, pb.table))
    lam_prev, _ = orc.adjoint_step(U, 15, dt, lam)
    rhs = np.sum(lam_prev * dU)
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # a block of a hanging slot takes part: perturbing it breaks the identity
    c, k = np.argwhere(np.abs(J[:, 5:]).sum(axis=(2, 3)) > 0)[0]
    J2 = J.copy()
    J2[c, 5 + k, 0, 0] += 1.0
    assert abs(np.sum(lam * _tangent(J2, dU, pb.table)) - rhs) > 1e-6 * abs(rhs)


def test_functional_gradient_matches_central_fd_on_adaptive_mesh(oracle_lib):
    pb = small_leaf_flow_problem(nsteps=25)
    orc = _leaf(oracle_lib, pb)
    box = (130.0, 260.0, 60.0, 180.0)
    w = pb.box_weights(box)
    UN, dts, lam0, _ = orc.forward_adjoint(_U(pb), pb.nsteps, w)
    J0 = np.sum(w.reshape(3, -1) * np.array(UN))
    assert J0 > 0
    t = pb.table
    rng = np.random.default_rng(2)
    wet = np.flatnonzero(pb.h > 0.5)
    fine = [c for c in wet if t.gen[c] == 2]
    checked = agree = 0
    for c in rng.choice(fine, size=12, replace=False):
        for comp in (0, 1):
            e = 1e-6 * max(abs(pb.h[c]), 1.0)
            vals = []
            for sgn in (+1, -1):
                U = _U(pb)
                U[comp][c] += sgn * e
                orc.time = 0.0
                orc.run(U, pb.nsteps, dt_seq=dts)
                vals.append(np.sum(w.reshape(3, -1) * np.array(U)))
            fd = (vals[0] - vals[1]) / (2 * e)
            checked += 1
            if abs(fd - lam0[comp, c]) <= 1e-5 * max(abs(fd), 1e-3 * abs(J0)):
                agree += 1
    assert agree >= checked - 2, (agree, checked)


def test_bed_friction_gradient_matches_central_fd_on_adaptive_mesh(oracle_lib):
    """J(phi_bed) of this run jumps (a switch flips) between -1e-5 and -1e-6
    degrees, so the difference is taken at 1e-7 degrees, inside the smooth
    piece; the one-sided difference at +1e-4 degrees agrees as well."""
    pb = small_leaf_flow_problem(nsteps=25)
    orc = _leaf(oracle_lib, pb)
    w = pb.box_weights((130.0, 260.0, 60.0, 180.0))
    _, dts, _, dphi = orc.forward_adjoint(_U(pb), pb.nsteps, w, want_dphi=True)

    def J(d):
        orc.set_angles(pb.phi_int_deg, pb.phi_bed_deg + d)
        U = _U(pb)
        orc.time = 0.0
        orc.run(U, pb.nsteps, dt_seq=dts)
        return np.sum(w.reshape(3, -1) * np.array(U))

    fd = (J(1e-7) - J(-1e-7)) / (2 * math.radians(1e-7))
    fwd = (J(1e-4) - J(0.0)) / math.radians(1e-4)
    orc.set_angles(pb.phi_int_deg, pb.phi_bed_deg)
    assert dphi != 0.0
    assert dphi == pytest.approx(fd, rel=1e-5)
    assert dphi == pytest.approx(fwd, rel=1e-5)
