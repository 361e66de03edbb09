"""GPU: the bed-friction gradient at the C2 size -- 512 x 512 cells (5 m) on
a seeded fractal DEM with one pile, 1000 forward steps, FP64 checkpoints
every 50, functional = deposit h over a seeded 20 x 20-cell target region --
against central finite differences of the functional in phi_bed
(eps = 1e-4 deg) with the recorded dt sequence frozen."""
import math

import numpy as np
import pytest

from discrete_adjoint_amd import problems
from discrete_adjoint_amd.solver import Simulation

pytestmark = pytest.mark.gpu

N, STEPS, EPS_DEG = 512, 1000, 1e-4


def test_dphi_matches_central_fd_at_512():
    pb = problems.fractal_problem(n=N, dx=5.0, nsteps=STEPS, npiles=1, seed=7, dem_seed=3,
                                  checkpoint_every=50)
    print(f"case C2 {N}x{N} {STEPS} steps", flush=True)
    with Simulation(pb, checkpoint_every=50, max_steps=STEPS) as sim:
        sim.run(STEPS)
        h = sim.get_state()[0]
        dts = sim.dt_record()
        box = problems.target_box(pb, h, cells=20, seed=5)
        w = pb.box_weights(box)
        J0 = float(np.sum(w[:pb.n] * h))
        _, dphi = sim.dual_solver(w, want_dphi=True, want_lambda=False)
    J = {}
    for sgn in (1, -1):
        with Simulation(pb, checkpoint_every=50, max_steps=STEPS) as sim:
            sim.set_params(pb.phi_int_deg, pb.phi_bed_deg + sgn * EPS_DEG)
            sim.replay(dts)
            J[sgn] = float(np.sum(w[:pb.n] * sim.get_state()[0]))
    fd = (J[1] - J[-1]) / (2 * math.radians(EPS_DEG))
    print(f"J {J0:.10g} adjoint dJ/dphi_bed {dphi:.8g} central FD {fd:.8g} "
          f"rel {abs(dphi - fd) / abs(fd):.3g}", flush=True)
    assert J0 > 0 and abs(fd) > 0
    assert abs(dphi - fd) <= 1e-3 * abs(fd)
