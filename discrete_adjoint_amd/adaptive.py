"""The time loop with mesh adaptation and its reverse sweep (hpfem.C:222-300
with H_adapt + unrefine every `adapt_every` steps; dual_solver.C:183-257
replaying the meshes backwards, dual_refine_unrefine.C:20-450).

Every adaptation pass is host code (tg_host_adapt); between passes the mesh
is fixed and runs on the GPU leaf path (tg_set_mesh / tg_forward_run /
tg_adjoint_run on one context).  A pass happens before step s when
s % adapt_every == adapt_every - 1 (TimeProps::ifrefine, properties.h:482-486),
so the run is a chain of fixed-mesh segments.  The reverse sweep takes the
segments last to first: it restores the segment's first state, replays its
recorded dt, runs the segment's adjoint from the adjoint its successor handed
back, and moves the result across the segment's adaptation pass with the
transposed transfer (tg_host_adapt_adjoint).

Trees are kept in tree order (the order tg_host_adapt returns); the GPU works
in the flattened element-table order of each mesh (tg_host_flatten).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib, mesh
from .problems import LeafProblem, _dem
from .solver import Simulation


def dem_statics(dem_args, length_scale=1.0, gravity_scale=1.0):
    """statics(xc) -> (grav (3, n), curv (2, n)) of leaf centres xc (n, 2)
    from a raster (nrows, ncols, xmin, ymax, res, elev float32):
    tg_host_leaf_static_coefs."""
    lib = _lib.load()
    dem = _dem(*dem_args)

    def statics(xc):
        x = np.ascontiguousarray(xc, dtype=np.float64).reshape(-1)
        n = len(x) // 2
        g, cv = np.empty(3 * n), np.empty(2 * n)
        rc = lib.tg_host_leaf_static_coefs(C.byref(dem), n, _lib.dptr(x), length_scale,
                                           gravity_scale, _lib.dptr(g), _lib.dptr(cv))
        if rc:
            raise ValueError(f"tg_host_leaf_static_coefs returned {rc}")
        return g.reshape(3, n), cv.reshape(2, n)
    statics._dem = dem  # keeps the raster alive
    return statics


@dataclass
class Segment:
    tree: tuple            # (lev, ix, iy) of the segment's mesh, tree order
    problem: LeafProblem   # flattened mesh and U at the segment's first step
    step0: int             # first step of the segment
    time0: float
    dts: np.ndarray        # dt of each step of the segment
    adapt: object = None   # the pass before the segment (mesh.AdaptPass) or None


@dataclass
class AdaptiveRun:
    """The adaptive time loop of one problem.  geometry: (nx0, ny0, xmin,
    ymin, dx0, dy0) of the base grid; tree and U (3, n) in tree order;
    statics(xc) -> (grav, curv) of leaf centres."""
    geometry: tuple
    tree: tuple
    U: np.ndarray
    statics: object
    phi_int_deg: float
    phi_bed_deg: float
    max_level: int
    adapt_every: int = 10
    num_buffer_layer: int = 5
    tiny: float = 1e-4
    checkpoint_every: int = 25
    use_graph: int = 1
    segments: list = field(default_factory=list)

    def problem_of(self, tree, U):
        """The flattened leaf mesh of a tree with the state U (tree order)."""
        nx0, ny0, xmin, ymin, dx0, dy0 = self.geometry
        lev, ix, iy = tree
        t = mesh.flatten(nx0, ny0, xmin, ymin, dx0, dy0, lev, ix, iy)
        grav, curv = self.statics(t.xc)
        dgrav = mesh.leaf_d_gravity(t, grav[2])
        o = t.order
        Uf = np.ascontiguousarray(np.asarray(U)[:, o])
        return LeafProblem("adaptive", nx0, ny0, dx0, dy0, xmin, ymin, t, grav, dgrav, curv,
                           Uf[0].copy(), Uf[1].copy(), Uf[2].copy(), self.phi_int_deg,
                           self.phi_bed_deg, self.adapt_every, self.checkpoint_every)

    def _bounds(self, nsteps):
        """Segment boundaries: 0, then every step s < nsteps with
        s % adapt_every == adapt_every - 1."""
        R = self.adapt_every
        cuts = [s for s in range(1, nsteps) if s % R == R - 1]
        return list(zip([0] + cuts, cuts + [nsteps]))

    def forward(self, nsteps):
        """Runs nsteps; returns the final tree and state (tree order).  The
        segments are kept for the reverse sweep."""
        self.close()
        self.segments = []
        tree, U = self.tree, np.array(self.U, dtype=np.float64, copy=True)
        time = 0.0
        sim = None
        for a, b in self._bounds(nsteps):
            adapt = None
            if a > 0:
                lev, ix, iy, Un, adapt = mesh.adapt(*self.geometry[:2], *tree, U,
                                                    self.max_level, self.num_buffer_layer,
                                                    self.tiny)
                tree, U = (lev, ix, iy), np.array(Un)
            pb = self.problem_of(tree, U)
            if sim is None:
                sim = Simulation(pb, checkpoint_every=self.checkpoint_every,
                                 max_steps=self.adapt_every, use_graph=self.use_graph)
            else:
                sim.set_problem(pb)
            sim.set_clock(a, time)
            time0 = time
            st = sim.run(b - a)
            time = st.time
            Uf = np.array(sim.get_state())
            U = np.empty_like(Uf)
            U[:, pb.table.order] = Uf
            self.segments.append(Segment(tree, pb, a, time0, sim.dt_record(), adapt))
        self._sim = sim
        self.final = (tree, U)
        return tree, U

    def adjoint(self, weights, want_dphi=True):
        """Reverse sweep of the functional J = sum weights . U^N (weights
        (3, n) on the final tree, tree order): returns (lambda^0 (3, n0) on the
        initial tree, dJ/dphi_bed)."""
        sim = self._sim
        lam = np.array(weights, dtype=np.float64).reshape(3, -1)
        dphi = 0.0
        for seg in reversed(self.segments):
            pb = seg.problem
            o = pb.table.order
            sim.set_problem(pb)
            sim.set_clock(seg.step0, seg.time0)
            sim.replay(seg.dts)
            lf, g = sim.dual_solver(lam[:, o], want_dphi=want_dphi)
            if want_dphi:
                dphi += g
            lam = np.empty_like(lf)
            lam[:, o] = lf
            if seg.adapt is not None:
                lam = seg.adapt.adjoint(lam)
        return lam, dphi

    def close(self):
        if getattr(self, "_sim", None) is not None:
            self._sim.close()
            self._sim = None
