"""TEST INFRASTRUCTURE ONLY: ctypes binding of the FP64 CPU restatement
(oracle/titan_oracle.c).  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; never by the product.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
OMP_LIB_PATH = os.path.join(HERE, "liboracle_omp.so")
_dp = C.POINTER(C.c_double)


class or_materials(C.Structure):
    _fields_ = [("mat", C.POINTER(C.c_ubyte)), ("tan_bed", _dp), ("kact_root", _dp),
                ("dkroot", _dp)]


class or_mesh(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double),
                ("gx", _dp), ("gy", _dp), ("gz", _dp), ("dgx", _dp), ("dgy", _dp),
                ("kx", _dp), ("ky", _dp), ("mats", or_materials)]


class or_params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "sin_int", "tan_bed", "kact_c1", "kact_root", "dkroot_dphi", "epsilon", "gscale",
        "tiny", "gamma", "cfl_inner", "cfl_outer", "maxtime", "nextoutput", "nextsave")]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_LIBS = {}


def load(omp=False):
    """The checker (omp False) or its OpenMP build (the all-cores CPU
    baseline: same per-cell results, FP64 sums in another order)."""
    path = OMP_LIB_PATH if omp else LIB_PATH
    if path not in _LIBS:
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        lib.or_step.restype = C.c_int
        lib.or_step.argtypes = [C.POINTER(or_mesh), C.POINTER(or_params), C.c_double, _dp, _dp,
                                _dp, C.c_double, _dp, _dp, _dp]
        lib.or_compute_dt.restype = C.c_double
        lib.or_compute_dt.argtypes = [C.POINTER(or_mesh), C.POINTER(or_params), _dp, _dp, _dp,
                                      C.c_double]
        lib.or_jacobian.restype = None
        lib.or_jacobian.argtypes = [C.POINTER(or_mesh), C.POINTER(or_params), C.c_double, _dp,
                                    _dp, _dp, C.c_double, _dp]
        lib.or_adjoint_step.restype = None
        lib.or_adjoint_step.argtypes = [C.POINTER(or_mesh), C.POINTER(or_params), C.c_double,
                                        _dp, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp]
        lib.or_dstep_dphi.restype = None
        lib.or_dstep_dphi.argtypes = [C.POINTER(or_mesh), C.POINTER(or_params), C.c_double, _dp,
                                      _dp, _dp, C.c_double, _dp]
        lib.or_switches.restype = None
        lib.or_switches.argtypes = [C.POINTER(or_mesh), C.POINTER(or_params), C.c_double, _dp,
                                    _dp, _dp, C.c_double, _dp, C.POINTER(C.c_int)]
        lib.or_kact_slopes.restype = None
        lib.or_kact_slopes.argtypes = [C.POINTER(or_mesh), C.POINTER(or_params), _dp, _dp, _dp,
                                       _dp, _dp]
        _LIBS[path] = lib
    return _LIBS[path]


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def bed_constants(phi_int_deg, phi_bed_deg):
    """MatProps + getcoef.f:39-43 constants of one bed friction angle, with
    libm: (tan_bed, kact_c1, kact_root, d kact_root / d phi_bed)."""
    pi_, pb_ = math.radians(phi_int_deg), math.radians(phi_bed_deg)
    cosphi, tandel = math.cos(pi_), math.tan(pb_)
    w = 1.0 - (1.0 + tandel * tandel) * (cosphi * cosphi)
    root = math.sqrt(abs(w))
    droot = (-(-1.0 if w < 0 else 1.0) * (cosphi * cosphi) * tandel * (1.0 + tandel * tandel) /
             root) if root > 0 else 0.0
    return tandel, 2.0 / (cosphi * cosphi), root, droot


def _materials(obj, mat, materials, phi_int_deg):
    """or_materials of per-cell indices `mat` and bed angles `materials`
    (kept alive on obj)."""
    mm = or_materials()
    if mat is None:
        return mm
    obj._mat = np.ascontiguousarray(mat, dtype=np.uint8).reshape(-1)
    consts = [bed_constants(phi_int_deg, b) for b in materials]
    obj._mtab = [np.array([c[i] for c in consts], dtype=np.float64) for i in (0, 2, 3)]
    mm.mat = obj._mat.ctypes.data_as(C.POINTER(C.c_ubyte))
    mm.tan_bed, mm.kact_root, mm.dkroot = [_p(a) for a in obj._mtab]
    return mm


def frict_tiny(s, switch=50, early=0.1, late=0.000000001):
    """hpfem.C:232-235 schedule for step index s (iter before the step)."""
    return early if s < switch else late


class Oracle:
    """FP64 CPU restatement bound to one uniform problem."""

    def __init__(self, nx, ny, dx, dy, x0, y0, grav, dgrav, curv, phi_int_deg, phi_bed_deg,
                 epsilon=1.0, gscale=1.0, tiny=1e-4, gamma=1.0, omp=False, mat=None,
                 materials=None):
        self.lib = load(omp)
        self.n = nx * ny
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
                      (grav[0], grav[1], grav[2], dgrav[0], dgrav[1], curv[0], curv[1])]
        m = or_mesh()
        m.nx, m.ny, m.dx, m.dy, m.x0, m.y0 = nx, ny, dx, dy, x0, y0
        m.gx, m.gy, m.gz, m.dgx, m.dgy, m.kx, m.ky = [_p(a) for a in self._keep]
        m.mats = _materials(self, mat, materials, phi_int_deg)
        self.mesh = m
        self.tiny, self.gamma, self.eps, self.gscale = tiny, gamma, epsilon, gscale
        self.set_angles(phi_int_deg, phi_bed_deg)
        self.time = 0.0

    @classmethod
    def from_problem(cls, pb, omp=False):
        return cls(pb.nx, pb.ny, pb.dx, pb.dy, pb.x0, pb.y0, pb.grav, pb.dgrav, pb.curv,
                   pb.phi_int_deg, pb.phi_bed_deg, pb.epsilon, pb.gravity_scale, omp=omp,
                   mat=getattr(pb, "mat", None), materials=getattr(pb, "materials", None))

    def set_angles(self, phi_int_deg, phi_bed_deg):
        """MatProps + getcoef.f:39-43 constants, evaluated with libm."""
        p = or_params()
        p.sin_int = math.sin(math.radians(phi_int_deg))
        p.tan_bed, p.kact_c1, p.kact_root, p.dkroot_dphi = bed_constants(phi_int_deg,
                                                                         phi_bed_deg)
        p.epsilon, p.gscale, p.tiny, p.gamma = self.eps, self.gscale, self.tiny, self.gamma
        p.cfl_inner = p.cfl_outer = 0.5
        p.maxtime = p.nextoutput = p.nextsave = math.inf
        self.params = p

    def step(self, U, s, dt=None):
        """One forward step in place on U = (h, hu, hv) (float64 arrays);
        returns (dt, mass)."""
        h, hu, hv = U
        t = C.c_double(self.time)
        dto, mass = C.c_double(), C.c_double()
        rc = self.lib.or_step(C.byref(self.mesh), C.byref(self.params), frict_tiny(s), _p(h),
                              _p(hu), _p(hv), math.nan if dt is None else float(dt), C.byref(t),
                              C.byref(dto), C.byref(mass))
        if rc:
            raise FloatingPointError("oracle: non-finite state")
        self.time = t.value
        return dto.value, mass.value

    def run(self, U, nsteps, dt_seq=None):
        dts = []
        for s in range(nsteps):
            dt, _ = self.step(U, s, None if dt_seq is None else dt_seq[s])
            dts.append(dt)
        return np.array(dts)

    def compute_dt(self, U):
        return self.lib.or_compute_dt(C.byref(self.mesh), C.byref(self.params), _p(U[0]),
                                      _p(U[1]), _p(U[2]), self.time)

    def kact_slopes(self, U):
        """(kact [n], slopes [n, 6]) of the state, as a forward step computes
        them before its fluxes."""
        U = [np.ascontiguousarray(a, dtype=np.float64) for a in U]
        k = np.empty(self.n)
        d = np.empty(6 * self.n)
        self.lib.or_kact_slopes(C.byref(self.mesh), C.byref(self.params), _p(U[0]), _p(U[1]),
                                _p(U[2]), _p(k), _p(d))
        return k, d.reshape(self.n, 6)

    def jacobian(self, U, s, dt):
        J = np.empty(self.n * 81)
        self.lib.or_jacobian(C.byref(self.mesh), C.byref(self.params), frict_tiny(s), _p(U[0]),
                             _p(U[1]), _p(U[2]), dt, _p(J))
        return J.reshape(self.n, 9, 3, 3)

    def adjoint_step(self, U, s, dt, lam_next, g=None, want_dphi=False):
        out = np.empty(3 * self.n)
        ln = np.ascontiguousarray(lam_next, dtype=np.float64).reshape(-1)
        gg = None if g is None else np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
        acc = C.c_double(0.0)
        self.lib.or_adjoint_step(C.byref(self.mesh), C.byref(self.params), frict_tiny(s),
                                 _p(U[0]), _p(U[1]), _p(U[2]), dt, _p(ln), _p(gg), _p(out),
                                 C.cast(C.byref(acc), _dp) if want_dphi else None)
        return out.reshape(3, -1), acc.value

    def error_update(self, bp, s, dt):
        """The state update_states predicts on an error grid (this oracle's
        mesh) from the reconstructed previous state bp (3, n), step s, dt."""
        b = [np.ascontiguousarray(x, dtype=np.float64) for x in bp]
        out = np.empty((3, self.n))
        self.lib.orl_error_update(C.byref(self.mesh), C.byref(self.params), frict_tiny(s),
                                  _p(b[0]), _p(b[1]), _p(b[2]), dt, _p(out[0]), _p(out[1]),
                                  _p(out[2]))
        return out

    def switches(self, U, s, dt):
        """(sg (2, n), stop (2, n)): the h-slope limiter's sign factor in x
        and y and the Coulomb stop flags of state U at step s (or_switches)."""
        sg = np.empty(2 * self.n)
        stop = np.empty(2 * self.n, dtype=np.int32)
        self.lib.or_switches(C.byref(self.mesh), C.byref(self.params), frict_tiny(s), _p(U[0]),
                             _p(U[1]), _p(U[2]), dt, _p(sg),
                             stop.ctypes.data_as(C.POINTER(C.c_int)))
        return sg.reshape(2, -1), stop.reshape(2, -1)

    def dstep_dphi(self, U, s, dt):
        out = np.empty(3 * self.n)
        self.lib.or_dstep_dphi(C.byref(self.mesh), C.byref(self.params), frict_tiny(s),
                               _p(U[0]), _p(U[1]), _p(U[2]), dt, _p(out))
        return out.reshape(3, -1)

    def forward_adjoint(self, U0, nsteps, terminal_weight, dt_seq=None, want_dphi=False):
        """Full store-everything forward + reverse sweep (SolRec semantics).
        Returns (U_N, dt_seq, lambda0, dJ/dphi_bed)."""
        U = [np.array(a, dtype=np.float64, copy=True) for a in U0]
        traj = []
        dts = []
        for s in range(nsteps):
            traj.append([a.copy() for a in U])
            dt, _ = self.step(U, s, None if dt_seq is None else dt_seq[s])
            dts.append(dt)
        lam = np.ascontiguousarray(terminal_weight, dtype=np.float64).reshape(3, -1)
        dphi = 0.0
        for s in range(nsteps - 1, -1, -1):
            lam, acc = self.adjoint_step(traj[s], s, dts[s], lam, want_dphi=want_dphi)
            dphi += acc
        return U, np.array(dts), lam, dphi


_ip = C.POINTER(C.c_int)


class or_leaf_mesh(C.Structure):
    _fields_ = [("n", C.c_int), ("nbr", _ip), ("nproc", _ip), ("gen", _ip), ("cdx", _dp),
                ("xc", _dp), ("gx", _dp), ("gy", _dp), ("gz", _dp), ("dgx", _dp), ("dgy", _dp),
                ("kx", _dp), ("ky", _dp), ("mats", or_materials)]


def _leaf_sigs(lib):
    M, P = C.POINTER(or_leaf_mesh), C.POINTER(or_params)
    lib.orl_step.restype = C.c_int
    lib.orl_step.argtypes = [M, P, C.c_double, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp]
    lib.orl_jacobian.restype = None
    lib.orl_jacobian.argtypes = [M, P, C.c_double, _dp, _dp, _dp, C.c_double, _dp]
    lib.orl_adjoint_step.restype = None
    lib.orl_adjoint_step.argtypes = [M, P, C.c_double, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp,
                                     _dp]
    lib.orl_dstep_dphi.restype = None
    lib.orl_dstep_dphi.argtypes = [M, P, C.c_double, _dp, _dp, _dp, C.c_double, _dp]
    lib.orl_slopes.restype = None
    lib.orl_slopes.argtypes = [M, P, _dp, _dp, _dp, _dp]
    lib.orl_error_update.restype = None
    lib.orl_error_update.argtypes = [M, P, C.c_double, _dp, _dp, _dp, C.c_double, _dp, _dp,
                                     _dp]


class LeafOracle(Oracle):
    """FP64 CPU restatement bound to one adaptive-mesh problem (leaf tables:
    nbr / nproc (n, 8), gen (n,), cdx / xc (n, 2))."""

    def __init__(self, nbr, nproc, gen, cdx, xc, grav, dgrav, curv, phi_int_deg, phi_bed_deg,
                 epsilon=1.0, gscale=1.0, tiny=1e-4, gamma=1.0, mat=None, materials=None):
        self.lib = load()
        _leaf_sigs(self.lib)
        self.n = len(gen)
        ints = [np.ascontiguousarray(a, dtype=np.intc).reshape(-1) for a in (nbr, nproc, gen)]
        dbl = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in
               (cdx, xc, grav[0], grav[1], grav[2], dgrav[0], dgrav[1], curv[0], curv[1])]
        self._keep = ints + dbl
        m = or_leaf_mesh()
        m.n = self.n
        m.nbr, m.nproc, m.gen = [a.ctypes.data_as(_ip) for a in ints]
        (m.cdx, m.xc, m.gx, m.gy, m.gz, m.dgx, m.dgy, m.kx, m.ky) = [_p(a) for a in dbl]
        m.mats = _materials(self, mat, materials, phi_int_deg)
        self.mesh = m
        self.tiny, self.gamma, self.eps, self.gscale = tiny, gamma, epsilon, gscale
        self.set_angles(phi_int_deg, phi_bed_deg)
        self.time = 0.0

    @classmethod
    def from_problem(cls, pb):
        t = pb.table
        return cls(t.nbr, t.nkind, t.gen, t.dx, t.xc, pb.grav, pb.dgrav, pb.curv,
                   pb.phi_int_deg, pb.phi_bed_deg, pb.epsilon, pb.gravity_scale,
                   mat=getattr(pb, "mat", None), materials=getattr(pb, "materials", None))

    def step(self, U, s, dt=None):
        h, hu, hv = U
        t = C.c_double(self.time)
        dto, mass = C.c_double(), C.c_double()
        rc = self.lib.orl_step(C.byref(self.mesh), C.byref(self.params), frict_tiny(s), _p(h),
                               _p(hu), _p(hv), math.nan if dt is None else float(dt), C.byref(t),
                               C.byref(dto), C.byref(mass))
        if rc:
            raise FloatingPointError("oracle: non-finite state")
        self.time = t.value
        return dto.value, mass.value

    def compute_dt(self, U):
        raise NotImplementedError

    def slopes(self, U):
        d = np.empty(6 * self.n)
        self.lib.orl_slopes(C.byref(self.mesh), C.byref(self.params), _p(U[0]), _p(U[1]),
                            _p(U[2]), _p(d))
        return d.reshape(self.n, 6)

    def jacobian(self, U, s, dt):
        J = np.empty(self.n * 81)
        self.lib.orl_jacobian(C.byref(self.mesh), C.byref(self.params), frict_tiny(s),
                              _p(U[0]), _p(U[1]), _p(U[2]), dt, _p(J))
        return J.reshape(self.n, 9, 3, 3)

    def adjoint_step(self, U, s, dt, lam_next, g=None, want_dphi=False):
        out = np.empty(3 * self.n)
        ln = np.ascontiguousarray(lam_next, dtype=np.float64).reshape(-1)
        gg = None if g is None else np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
        acc = C.c_double(0.0)
        self.lib.orl_adjoint_step(C.byref(self.mesh), C.byref(self.params), frict_tiny(s),
                                  _p(U[0]), _p(U[1]), _p(U[2]), dt, _p(ln), _p(gg), _p(out),
                                  C.cast(C.byref(acc), _dp) if want_dphi else None)
        return out.reshape(3, -1), acc.value

    def error_update(self, bp, s, dt):
        """The state update_states predicts on an error grid (this oracle's
        mesh) from the reconstructed previous state bp (3, n), step s, dt."""
        b = [np.ascontiguousarray(x, dtype=np.float64) for x in bp]
        out = np.empty((3, self.n))
        self.lib.orl_error_update(C.byref(self.mesh), C.byref(self.params), frict_tiny(s),
                                  _p(b[0]), _p(b[1]), _p(b[2]), dt, _p(out[0]), _p(out[1]),
                                  _p(out[2]))
        return out

    def switches(self, U, s, dt):
        """(sg (2, n), stop (2, n)): the h-slope limiter's sign factor in x
        and y and the Coulomb stop flags of state U at step s (or_switches)."""
        sg = np.empty(2 * self.n)
        stop = np.empty(2 * self.n, dtype=np.int32)
        self.lib.or_switches(C.byref(self.mesh), C.byref(self.params), frict_tiny(s), _p(U[0]),
                             _p(U[1]), _p(U[2]), dt, _p(sg),
                             stop.ctypes.data_as(C.POINTER(C.c_int)))
        return sg.reshape(2, -1), stop.reshape(2, -1)

    def dstep_dphi(self, U, s, dt):
        out = np.empty(3 * self.n)
        self.lib.orl_dstep_dphi(C.byref(self.mesh), C.byref(self.params), frict_tiny(s),
                                _p(U[0]), _p(U[1]), _p(U[2]), dt, _p(out))
        return out.reshape(3, -1)
