"""Pins the FP64 CPU restatement (oracle/) with checks that do not retype
its formulas: exact invariants, closed forms, finite differences of the
forward map against the analytic Jacobian blocks (the reference's
compare_jacobians, debug_routines.C:36-57), the adjoint dot-product
identity, and end-to-end finite differences of the functional and of the
bed-friction gradient at frozen dt.

The reference ships no tests, golden vectors or fixtures for this path, and
its build needs MPI/HDF5/LAPACK (absent), so these are the pins."""
import math

import numpy as np
import pytest

from helpers import analytic_problem, paraboloid, small_flow_problem


def _orc(oracle_lib, pb):
    return oracle_lib.Oracle.from_problem(pb)


def _U(pb):
    return [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]


def test_lake_at_rest_is_exact(oracle_lib):
    # flat ground, uniform depth, no motion: every face flux is the same
    # physical flux, sources vanish -> U unchanged to the last bit.
    pb = analytic_problem(16, 12, h=np.full(16 * 12, 2.5))
    orc = _orc(oracle_lib, pb)
    U = _U(pb)
    for s in range(5):
        orc.step(U, s)
    assert np.array_equal(U[0], pb.h) and not U[1].any() and not U[2].any()


def test_cfl_closed_form_at_rest(oracle_lib):
    # h = 1 at rest on flat ground: kact = eps (no yielding, getcoef.f:48-51),
    # evalue = sqrt(eps*g*h); dt = 0.5*min(0.5*dx/evalue, 0.9*sqrt(hmax*eps/9.8))
    pb = analytic_problem(8, 8, dx=10.0, h=np.ones(64))
    orc = _orc(oracle_lib, pb)
    dt = orc.compute_dt(_U(pb))
    ev = math.sqrt(1.0 * 9.8 * 1.0)
    assert dt == pytest.approx(0.5 * min(0.5 * 10.0 / ev, 0.9 * math.sqrt(1.0 / 9.8)), rel=1e-15)
    # thin: the height cap wins
    pb = analytic_problem(8, 8, dx=10.0, h=np.full(64, 0.01))
    dt = _orc(oracle_lib, pb).compute_dt(_U(pb))
    assert dt == pytest.approx(0.5 * 0.9 * math.sqrt(0.01 / 9.8), rel=1e-15)


def test_cfl_closed_form_moving_kact(oracle_lib):
    # uniform h=1, hu=2: zero slopes -> vel = 0 -> sgn = 0 ->
    # K = 2/cos^2(phi_int) - 1; evalue = |u| + sqrt(K g h)
    phi = 30.0
    pb = analytic_problem(8, 8, dx=10.0, h=np.ones(64), hu=np.full(64, 2.0), phi_int=phi)
    dt = _orc(oracle_lib, pb).compute_dt(_U(pb))
    K = 2.0 / math.cos(math.radians(phi)) ** 2 - 1.0
    ev = 2.0 + math.sqrt(K * 9.8)
    assert dt == pytest.approx(0.5 * min(0.5 * 10.0 / ev, 0.9 * math.sqrt(1.0 / 9.8)), rel=1e-14)


def test_uniform_sheet_frictionless_gravity(oracle_lib):
    # a uniform sheet on a frictionless incline accelerates at g_x exactly
    # (uniform state -> no flux divergence, internal friction sign 0)
    n = 10 * 10
    pb = analytic_problem(10, 10, slope_x_deg=20.0, h=np.full(n, 3.0), phi_bed=0.0,
                          phi_int=30.0)
    orc = _orc(oracle_lib, pb)
    U = _U(pb)
    hu = 0.0
    for s in range(6):
        dt, _ = orc.step(U, s)
        hu = hu + dt * (pb.grav[0, 0] * 3.0 - 0.0 - 0.0)
        assert np.all(U[1] == hu) and np.all(U[0] == 3.0)
    assert hu > 0


def test_mass_conservation_and_positivity(oracle_lib):
    # boundary cells copy the inner face flux onto the outer face, so they
    # swallow whatever enters them (outflow); mass is conserved exactly
    # while the first interior ring stays dry.
    pb = small_flow_problem(nx=40, ny=30, nsteps=40)
    orc = _orc(oracle_lib, pb)
    U = _U(pb)
    m0 = pb.h.sum() * pb.dx * pb.dy
    for s in range(pb.nsteps):
        _, m = orc.step(U, s)
        assert U[0].min() >= 0.0
    hh = U[0].reshape(pb.ny, pb.nx)
    assert hh[:2].max() == 0 and hh[-2:].max() == 0 and hh[:, :2].max() == 0
    assert hh[:, -2:].max() == 0
    assert m == pytest.approx(m0, rel=1e-13)
    # the outflow cells do swallow mass once the front reaches them
    pb = small_flow_problem(nx=24, ny=20, nsteps=40)
    orc = _orc(oracle_lib, pb)
    U = _U(pb)
    for s in range(pb.nsteps):
        _, m = orc.step(U, s)
    assert m < pb.h.sum() * pb.dx * pb.dy * (1 - 1e-5)


def test_mirror_symmetry(oracle_lib):
    # flat ground, pile symmetric about the vertical mid-line: h is mirror
    # symmetric and hu antisymmetric to the last bit (x -> -x)
    nx, ny = 20, 16
    pb = analytic_problem(nx, ny, h=paraboloid(100.0, 70.0, 60.0, 40.0, 5.0))
    orc = _orc(oracle_lib, pb)
    U = _U(pb)
    for s in range(25):
        orc.step(U, s)
    h = U[0].reshape(ny, nx)
    hu = U[1].reshape(ny, nx)
    hv = U[2].reshape(ny, nx)
    assert np.array_equal(h, h[:, ::-1])
    assert np.array_equal(hu, -hu[:, ::-1])
    assert np.array_equal(hv, hv[:, ::-1])
    assert np.abs(hu).max() > 0


def _advance(orc, U, nsteps):
    dts = []
    for s in range(nsteps):
        dt, _ = orc.step(U, s)
        dts.append(dt)
    return dts


def _fd_step(orc, U, s, dt, c, comp, eps):
    outs = []
    for sgn in (+1, -1):
        V = [a.copy() for a in U]
        V[comp][c] += sgn * eps
        t = orc.time
        orc.step(V, s, dt=dt)
        orc.time = t
        outs.append(np.array(V))
    return (outs[0] - outs[1]) / (2 * eps)


@pytest.mark.parametrize("s_extra", [0, 55])
def test_jacobian_blocks_match_finite_differences(oracle_lib, s_extra):
    """calc_jacobian's 3x3 blocks equal -dU^{s+1}/dU^s from central FD of
    the forward step (frozen dt), where no switch flips in the FD stencil
    (the reference's comparator compares analytic vs FD the same way)."""
    pb = small_flow_problem(nsteps=20)
    orc = _orc(oracle_lib, pb)
    U = _U(pb)
    _advance(orc, U, 20)
    s = 20 + s_extra            # 20 -> frict_tiny 0.1, 75 -> 1e-9
    dt = orc.compute_dt(U)
    J = orc.jacobian(U, s, dt)
    nx = pb.nx
    rng = np.random.default_rng(0)
    wet = np.flatnonzero(U[0] > 0.05)
    wet = [c for c in wet if 1 < c % nx < nx - 2 and 1 < c // nx < pb.ny - 2]
    checked = agree = 0
    for c in rng.choice(wet, size=min(40, len(wet)), replace=False):
        for comp in range(3):
            scale = max(abs(U[comp][c]), 1e-2)
            e1 = 1e-6 * scale
            fd1 = _fd_step(orc, U, s, dt, c, comp, e1)
            fd2 = _fd_step(orc, U, s, dt, c, comp, 2 * e1)
            # cells j whose residual depends on U_c: c itself and 4 neighbours
            # (slot of c in j's blocks: self=0, E=1, N=2, W=3, S=4 in oracle order)
            for j, slot in ((c, 0), (c - 1, 1), (c - nx, 2), (c + 1, 3), (c + nx, 4)):
                an = -J[j, slot][:, comp]
                f1, f2 = fd1[:, j], fd2[:, j]
                if not np.allclose(f1, f2, rtol=1e-4, atol=1e-9):
                    continue  # a switch flipped inside the stencil
                checked += 1
                if np.allclose(an, f1, rtol=1e-5, atol=1e-7 * max(1.0, np.abs(f1).max())):
                    agree += 1
    assert checked > 300
    assert agree >= 0.97 * checked, (agree, checked)


def _liftoff_problem():
    """A fast sheet over a convex bed (curvature -0.02 / m along x): the
    Coulomb term's max(g_z h + u hu k_x, 0) clamps (stop) on most wet cells."""
    return analytic_problem(30, 20, 10.0, slope_x_deg=30.0, slope_y_deg=2.0,
                            h=paraboloid(70.0, 100.0, 60.0, 50.0, 8.0), phi_int=32.0,
                            phi_bed=15.0, nsteps=40, curv=(-0.02, 0.0))


@pytest.mark.parametrize("case", ["flow", "liftoff"])
def test_jacobian_fd_mismatches_are_the_reference_approximations(oracle_lib, case):
    """Where the analytic blocks differ from central FD of the forward step
    (frozen dt, no switch flipping inside the FD stencil), the block belongs
    to a cell where one of the reference's own adjoint approximations acts:
    the h-slope limiter with one difference in the c_sgn dead zone (sign
    factor 1/2, which hslope_sens does not carry), or the Coulomb stop (the
    neighbour blocks drop the internal-friction slope term,
    calc_jacobain.C), or the positivity clamp of h^{s+1} (not
    differentiated).  Every other block of every interior wet cell equals
    FD, and both approximations do show up.  So the gap between the adjoint
    and FD derivatives (the C2 comparison) is the reference's, not a slip of
    the restatement."""
    if case == "flow":
        pb, nadv, s = small_flow_problem(nsteps=20), 20, 75  # frict_tiny 1e-9
    else:
        pb, nadv, s = _liftoff_problem(), 40, 40
    orc = _orc(oracle_lib, pb)
    U = _U(pb)
    _advance(orc, U, nadv)
    dt = orc.compute_dt(U)
    J = orc.jacobian(U, s, dt)
    sg, stop = orc.switches(U, s, dt)
    nxt = [a.copy() for a in U]
    t = orc.time
    orc.step(nxt, s, dt=dt)
    orc.time = t
    nx, ny = pb.nx, pb.ny
    cells = [c for c in range(pb.n)
             if 1 < c % nx < nx - 2 and 1 < c // nx < ny - 2 and U[0][c] > 1e-3]
    checked = 0
    why = {"half-sign": 0, "stop": 0, "clamp": 0}
    unexplained = []
    for c in cells:
        for comp in range(3):
            e1 = 1e-6 * max(abs(U[comp][c]), U[0][c], 1e-2)
            fd1 = _fd_step(orc, U, s, dt, c, comp, e1)
            fd2 = _fd_step(orc, U, s, dt, c, comp, 2 * e1)
            for j, slot in ((c, 0), (c - 1, 1), (c - nx, 2), (c + 1, 3), (c + nx, 4)):
                an = -J[j, slot][:, comp]
                f1, f2 = fd1[:, j], fd2[:, j]
                if not np.allclose(f1, f2, rtol=1e-4, atol=1e-9):
                    continue  # a switch flipped inside the stencil
                checked += 1
                if np.allclose(an, f1, rtol=1e-5, atol=1e-7 * max(1.0, np.abs(f1).max())):
                    continue
                if abs(sg[0, j]) == 0.5 or abs(sg[1, j]) == 0.5:
                    why["half-sign"] += 1
                elif stop[:, j].any():
                    why["stop"] += 1
                elif nxt[0][j] == 0.0:
                    why["clamp"] += 1
                else:
                    unexplained.append((j, slot, comp))
    print(case, "checked", checked, "mismatches by cause", why)
    assert checked > 3000
    assert not unexplained, unexplained[:10]
    assert why["half-sign" if case == "flow" else "stop"] > 0


def _tangent(J, dU, nx, ny):
    """delta U^{s+1}_j = -sum_slots J_j[slot] delta U^s_{nbr(j,slot)}."""
    n = nx * ny
    out = np.zeros((3, n))
    dUt = dU.reshape(3, n)
    idx = np.arange(n)
    i, j = idx % nx, idx // nx
    nb = [idx, np.where(i + 1 < nx, idx + 1, -1), np.where(j + 1 < ny, idx + nx, -1),
          np.where(i > 0, idx - 1, -1), np.where(j > 0, idx - nx, -1)]
    for slot in range(5):
        m = nb[slot] >= 0
        src = dUt[:, nb[slot][m]]              # (3, k)
        out[:, m] -= np.einsum("kab,bk->ak", J[m, slot], src)
    return out


def test_adjoint_dot_product_identity(oracle_lib):
    """<lambda^{s+1}, T dU> == <T^T lambda^{s+1}, dU> where T dU applies the
    stored blocks forward and T^T is calc_adjoint's gather (which_neighbor
    indexing, calc_adjoint.C:110-140)."""
    pb = small_flow_problem(nsteps=15)
    orc = _orc(oracle_lib, pb)
    U = _U(pb)
    _advance(orc, U, 15)
    dt = orc.compute_dt(U)
    J = orc.jacobian(U, 15, dt)
    rng = np.random.default_rng(1)
    dU = rng.standard_normal((3, pb.n))
    lam = rng.standard_normal((3, pb.n))
    lhs = np.sum(lam * _tangent(J, dU, pb.nx, pb.ny))
    lam_prev, _ = orc.adjoint_step(U, 15, dt, lam)
    rhs = np.sum(lam_prev * dU)
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # the gather sees every block: perturbing one block entry breaks it
    J2 = J.copy()
    J2[(pb.ny // 2) * pb.nx + pb.nx // 2, 3, 1, 0] += 1.0
    lhs2 = np.sum(lam * _tangent(J2, dU, pb.nx, pb.ny))
    assert abs(lhs2 - rhs) > 1e-6 * abs(rhs)


def test_functional_gradient_matches_central_fd(oracle_lib):
    """lambda^0 = dJ/dU^0 for J = sum_box h^N dx dy, against central FD of
    the whole run at frozen dt (the reference's perturbU test,
    adjoint_test.C:133-326)."""
    pb = small_flow_problem(nsteps=25)
    orc = _orc(oracle_lib, pb)
    nx = pb.nx
    box = (120.0, 200.0, 40.0, 160.0)
    w = pb.box_weights(box)
    UN, dts, lam0, _ = orc.forward_adjoint(_U(pb), pb.nsteps, w)
    J0 = np.sum(w.reshape(3, -1) * np.array(UN))
    assert J0 > 0
    rng = np.random.default_rng(2)
    wet = np.flatnonzero(pb.h > 0.5)
    checked = agree = 0
    for c in rng.choice(wet, size=12, replace=False):
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


def test_bed_friction_gradient_matches_central_fd(oracle_lib):
    """dJ/dphi_bed from the reverse sweep vs central FD with eps = 1e-4 deg
    at frozen dt (no reference counterpart; the derivation is checked here)."""
    pb = small_flow_problem(nsteps=25)
    orc = _orc(oracle_lib, pb)
    w = pb.box_weights((120.0, 200.0, 40.0, 160.0))
    _, dts, _, dphi = orc.forward_adjoint(_U(pb), pb.nsteps, w, want_dphi=True)
    vals = []
    for sgn in (+1, -1):
        orc.set_angles(pb.phi_int_deg, pb.phi_bed_deg + sgn * 1e-4)
        U = _U(pb)
        orc.time = 0.0
        orc.run(U, pb.nsteps, dt_seq=dts)
        vals.append(np.sum(w.reshape(3, -1) * np.array(U)))
    orc.set_angles(pb.phi_int_deg, pb.phi_bed_deg)
    fd = (vals[0] - vals[1]) / (2 * math.radians(1e-4))
    assert dphi != 0.0
    assert dphi == pytest.approx(fd, rel=2e-4)


def test_checkpoint_free_reverse_is_deterministic(oracle_lib):
    pb = small_flow_problem(nsteps=12)
    orc = _orc(oracle_lib, pb)
    w = pb.box_weights((120.0, 200.0, 40.0, 160.0))
    a = orc.forward_adjoint(_U(pb), 12, w)
    orc.time = 0.0
    b = orc.forward_adjoint(_U(pb), 12, w)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[1], b[1])


def test_planted_sign_error_is_caught(oracle_lib):
    """A dropped minus sign in the reverse gather (a plausible slip) breaks
    the FD agreement: guards the FD test's discriminating power."""
    pb = small_flow_problem(nsteps=10)
    orc = _orc(oracle_lib, pb)
    w = pb.box_weights((120.0, 200.0, 40.0, 160.0))
    _, dts, lam0, _ = orc.forward_adjoint(_U(pb), 10, w)
    c = int(np.argmax(pb.h))
    e = 1e-6 * pb.h[c]
    vals = []
    for sgn in (+1, -1):
        U = _U(pb)
        U[0][c] += sgn * e
        orc.time = 0.0
        orc.run(U, 10, dt_seq=dts)
        vals.append(np.sum(w.reshape(3, -1) * np.array(U)))
    fd = (vals[0] - vals[1]) / (2 * e)
    assert lam0[0, c] == pytest.approx(fd, rel=1e-5, abs=1e-8)
    assert not (-lam0[0, c] == pytest.approx(fd, rel=1e-5, abs=1e-8))


def test_openmp_build_matches_the_checker(oracle_lib):
    # the all-cores CPU baseline runs the OpenMP build of the same source:
    # per-cell results bit-identical, sums within rounding
    import os
    import subprocess
    import sys
    code = (
        "import numpy as np, oracle\n"
        "from helpers import small_flow_problem\n"
        "pb = small_flow_problem(40, 30, 30)\n"
        "w = np.zeros((3, pb.n)); w[0, pb.n // 2:] = 1.0\n"
        "out = []\n"
        "for omp in (False, True):\n"
        "    orc = oracle.Oracle.from_problem(pb, omp=omp)\n"
        "    UN, dts, lam, dphi = orc.forward_adjoint([pb.h, pb.hu, pb.hv], 30, w, want_dphi=True)\n"
        "    out.append((UN, dts, lam, dphi))\n"
        "(a, b) = out\n"
        "assert all(np.array_equal(x, y) for x, y in zip(a[0], b[0]))\n"
        "assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])\n"
        "assert abs(a[3]) > 1e-3 and abs(a[3] - b[3]) <= 1e-12 * abs(a[3])\n"
        "print('ok')\n")
    here = os.path.dirname(os.path.abspath(__file__))
    env = dict(os.environ, OMP_NUM_THREADS="4",
               PYTHONPATH=os.pathsep.join([os.path.dirname(here), here]))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr
