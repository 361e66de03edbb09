"""GPU: the bed-friction gradient at the C2 grid size -- 512 x 512 cells (5 m)
on a seeded fractal DEM with one pile, FP64 checkpoints every 50, functional
= deposit h over a seeded 20 x 20-cell target region -- against finite
differences of the functional in phi_bed with the recorded dt sequence
frozen.

J(phi_bed) of this scheme is piecewise smooth: wet/dry thresholds, the kact
active/passive switch and the Coulomb stop make it jump wherever a switch
flips inside the perturbation.  Over 1000 steps on 512^2 cells the jumps are
dense down to 1e-7 deg, so no difference quotient converges there (the
central difference at 1e-4 deg is printed, not asserted).  Over 200 steps
the 1e-7 deg difference straddles no jump and is the derivative of the
smooth piece, which the adjoint gives; the two agree to the accuracy of the
reference's own adjoint approximations (slope sensitivities in the limiter
dead zone, the neighbour friction term under `stop`), a few 1e-4.
"""
import math

import numpy as np
import pytest

from discrete_adjoint_amd import problems
from discrete_adjoint_amd.solver import Simulation

pytestmark = pytest.mark.gpu

N = 512


def _case(steps):
    pb = problems.fractal_problem(n=N, dx=5.0, nsteps=steps, npiles=1, seed=7, dem_seed=3,
                                  checkpoint_every=50)
    with Simulation(pb, checkpoint_every=50, max_steps=steps) as sim:
        sim.run(steps)
        h = sim.get_state()[0]
        dts = sim.dt_record()
        box = problems.target_box(pb, h, cells=20, seed=5)
        w = pb.box_weights(box)
        J0 = float(np.sum(w[:pb.n] * h))
        _, dphi = sim.dual_solver(w, want_dphi=True, want_lambda=False)

    def J(ddeg):
        with Simulation(pb, checkpoint_every=50, max_steps=steps) as sim:
            sim.set_params(pb.phi_int_deg, pb.phi_bed_deg + ddeg)
            sim.replay(dts)
            return float(np.sum(w[:pb.n] * sim.get_state()[0]))

    return J0, dphi, J


def test_dphi_matches_fd_of_the_smooth_piece_at_512():
    print(f"case C2 {N}x{N} 200 steps", flush=True)
    J0, dphi, J = _case(200)
    assert J(0.0) == J0  # the frozen-dt replay reproduces the run
    e = 1e-7
    fd = (J(e) - J(-e)) / (2 * math.radians(e))
    fd4 = (J(1e-4) - J(-1e-4)) / (2 * math.radians(1e-4))
    print(f"J {J0:.10g} adjoint {dphi:.8g} central FD 1e-7 deg {fd:.8g} "
          f"(rel {abs(dphi - fd) / abs(fd):.3g}); central FD 1e-4 deg {fd4:.8g}", flush=True)
    assert abs(dphi - fd) <= 1e-3 * abs(fd)


def test_c2_1000_steps_gradient_is_finite_and_replay_exact():
    print(f"case C2 {N}x{N} 1000 steps", flush=True)
    J0, dphi, J = _case(1000)
    assert J(0.0) == J0
    fd4 = (J(1e-4) - J(-1e-4)) / (2 * math.radians(1e-4))
    print(f"J {J0:.10g} adjoint {dphi:.8g} central FD 1e-4 deg {fd4:.8g}", flush=True)
    assert math.isfinite(dphi)
