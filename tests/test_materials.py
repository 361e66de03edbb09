"""Per-cell materials (Element::material; the bed friction read as
bedfrict[material] in get_coef_and_eigen.C:112-114, corrector.C:108 and the
adjoint's calc_jacobian): the oracle's lookup pinned on the CPU.

- every cell of a uniform sheet on a plane accelerates at
  g (sin theta - cos theta tan phi_bed) of its own material;
- two materials of equal angle give the one-material results bit for bit
  (uniform blocks and adaptive meshes);
- different angles change the flow;
- dJ/dphi_bed with materials is the derivative with respect to one common
  change of every material's angle: central FD at frozen dt.
"""
import math

import numpy as np
import pytest

import oracle
from helpers import small_flow_problem, small_leaf_flow_problem, with_materials


def _U(pb):
    return [pb.h.copy(), pb.hu.copy(), pb.hv.copy()]


def test_sheet_cells_follow_their_own_material(oracle_lib):
    nx, ny, dx, th = 10, 6, 10.0, math.radians(30.0)
    n = nx * ny
    grav = np.stack([np.full(n, 9.8 * math.sin(th)), np.zeros(n), np.full(n, 9.8 * math.cos(th))])
    mat = np.tile((np.arange(nx) >= 5).astype(np.uint8), ny)
    orc = oracle.Oracle(nx, ny, dx, dx, 5.0, 5.0, grav, np.zeros((2, n)), np.zeros((2, n)),
                        30.0, 99.0, mat=mat, materials=(15.0, 25.0))
    u0, dt = 2.0, 0.01
    U = [np.ones(n), np.full(n, u0), np.zeros(n)]
    orc.step(U, 60, dt)
    for m, phi in ((0, 15.0), (1, 25.0)):
        acc = 9.8 * (math.sin(th) - math.cos(th) * math.tan(math.radians(phi)))
        np.testing.assert_allclose(U[1][mat == m], u0 + dt * acc, rtol=1e-14)


@pytest.mark.parametrize("leaves", [False, True])
def test_equal_materials_equal_one_material(oracle_lib, leaves):
    pb = small_leaf_flow_problem(nsteps=15) if leaves else small_flow_problem(nsteps=15)
    q = with_materials(pb, (pb.phi_bed_deg, pb.phi_bed_deg))
    assert q.mat.any() and not q.mat.all()
    make = oracle.LeafOracle.from_problem if leaves else oracle.Oracle.from_problem
    w = np.zeros((3, pb.n))
    w[0, : pb.n // 2] = 1.0
    a = make(pb).forward_adjoint(_U(pb), 15, w, want_dphi=True)
    b = make(q).forward_adjoint(_U(q), 15, w, want_dphi=True)
    for x, y in zip(a[0], b[0]):
        assert np.array_equal(x, y)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]) and a[3] == b[3]


def test_materials_change_the_flow(oracle_lib):
    pb = small_flow_problem(nsteps=30)
    q = with_materials(pb, (pb.phi_bed_deg, 35.0), split=0.4)
    U1, U2 = _U(pb), _U(q)
    oracle.Oracle.from_problem(pb).run(U1, 30)
    oracle.Oracle.from_problem(q).run(U2, 30)
    sel = q.mat == 1
    assert not np.array_equal(U1[1][sel], U2[1][sel])
    # the high-friction material holds the flow back
    assert np.abs(U2[1][sel]).sum() < np.abs(U1[1][sel]).sum()


@pytest.mark.parametrize("leaves", [False, True])
def test_common_shift_gradient_matches_fd(oracle_lib, leaves):
    pb = small_leaf_flow_problem(nsteps=20) if leaves else small_flow_problem(nsteps=20)
    q = with_materials(pb, (pb.phi_bed_deg, pb.phi_bed_deg + 8.0), split=0.4)
    make = oracle.LeafOracle.from_problem if leaves else oracle.Oracle.from_problem
    w = q.box_weights((120.0, 200.0, 40.0, 160.0)) if not leaves else \
        q.box_weights((100.0, 220.0, 40.0, 200.0))
    _, dts, _, dphi = make(q).forward_adjoint(_U(q), 20, w, want_dphi=True)
    eps = 1e-4 if not leaves else 1e-7
    vals = []
    for sgn in (+1, -1):
        r = with_materials(pb, (pb.phi_bed_deg + sgn * eps, pb.phi_bed_deg + 8.0 + sgn * eps),
                           split=0.4)
        r.phi_bed_deg = pb.phi_bed_deg + sgn * eps
        U = _U(r)
        make(r).run(U, 20, dt_seq=dts)
        vals.append(np.sum(w.reshape(3, -1) * np.array(U)))
    fd = (vals[0] - vals[1]) / (2 * math.radians(eps))
    assert dphi != 0.0
    assert dphi == pytest.approx(fd, rel=2e-4)
