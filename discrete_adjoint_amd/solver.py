"""Host-side interface mirroring the reference's entry points for the hot
path, over the C ABI (include/tg.h):

  Simulation.step()        <- step()            src/geoflow/step.C:29
  Simulation.run(n)        <- hpfem.C:222-300 time loop (no adaptation)
  Simulation.dual_solver() <- dual_solver()     src/adjoint/dual_solver.C:80
  Simulation.adjoint_step  <- calc_jacobian + calc_adjoint of one reverse step

Errors: the reference prints and exits; here every non-zero tg_status raises
TgError carrying the library's message.
"""
import ctypes as C
import math

import numpy as np

from . import _lib


class TgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tg error {code}: {msg}")
        self.code = code


def default_config(**kw):
    lib = _lib.load()
    cfg = _lib.tg_config()
    lib.tg_config_default(C.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def params_from_angles(phi_int_deg, phi_bed_deg, epsilon=1.0, gravity_scale=1.0):
    lib = _lib.load()
    p = _lib.tg_params()
    rc = lib.tg_params_from_angles(math.radians(phi_int_deg), math.radians(phi_bed_deg),
                                   epsilon, gravity_scale, C.byref(p))
    if rc:
        raise TgError(rc, "tg_params_from_angles")
    return p


class Simulation:
    """One problem (a uniform block, or an adaptive mesh given as leaf
    tables) on one GPU."""

    def __init__(self, problem, precision=_lib.TG_FP64, checkpoint_every=None, max_steps=None,
                 device=0, **cfg_overrides):
        self.lib = _lib.load()
        self.problem = problem
        ck = checkpoint_every or problem.checkpoint_every
        ms = max_steps or max(problem.nsteps, 1)
        self.cfg = default_config(precision=precision, checkpoint_every=ck, max_steps=ms,
                                  device=device, **cfg_overrides)
        self.ctx = C.c_void_p()
        self._check(self.lib.tg_create(C.byref(self.cfg), C.byref(self.ctx)), nocontext=True)
        self.set_problem(problem)

    def set_problem(self, problem):
        """Hands a (new) mesh with its static coefficients, materials and
        initial state to the context (tg_set_mesh: a re-mesh keeps the
        context's configuration)."""
        self.problem = problem
        self._arrays = []
        m = _lib.tg_mesh()
        if hasattr(problem, "table"):  # adaptive mesh: the flattened leaf tables
            t = problem.table
            tabs = [np.ascontiguousarray(a, dtype=np.int32).reshape(-1)
                    for a in (t.nbr, t.nkind, t.gen)]
            geo = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (t.dx, t.xc)]
            self._arrays += tabs + geo
            m.kind = _lib.TG_MESH_LEAVES
            m.n = problem.n
            m.nbr, m.nproc, m.gen = [_lib.iptr(a) for a in tabs]
            m.cdx, m.xc = [_lib.dptr(a) for a in geo]
        else:
            m.kind = _lib.TG_MESH_UNIFORM
            m.nx, m.ny, m.dx, m.dy, m.x0, m.y0 = (problem.nx, problem.ny, problem.dx, problem.dy,
                                                  problem.x0, problem.y0)
        g = np.ascontiguousarray(problem.grav.reshape(-1))
        dg = np.ascontiguousarray(problem.dgrav.reshape(-1))
        cv = np.ascontiguousarray(problem.curv.reshape(-1))
        self._arrays += [g, dg, cv]
        m.grav, m.dgrav, m.curv = _lib.dptr(g), _lib.dptr(dg), _lib.dptr(cv)
        _mesh_materials(m, problem, self._arrays)
        self._check(self.lib.tg_set_mesh(self.ctx, C.byref(m)))
        self._set_materials(0, problem)
        self.set_params(problem.phi_int_deg, problem.phi_bed_deg)
        self.set_state(problem.h, problem.hu, problem.hv)

    def set_clock(self, step0, time):
        """The state is U^{step0} at simulated time `time` (tg_set_clock)."""
        self._check(self.lib.tg_set_clock(self.ctx, int(step0), float(time)))

    # -- plumbing -------------------------------------------------------------
    def _set_materials(self, member, problem):
        """tg_set_materials from the problem's per-material bed friction angles."""
        if problem.mat is None or not problem.materials:
            return
        mats = (_lib.tg_material * len(problem.materials))()
        for i, phi_bed in enumerate(problem.materials):
            rc = self.lib.tg_material_from_angles(math.radians(problem.phi_int_deg),
                                                  math.radians(phi_bed), C.byref(mats[i]))
            if rc:
                raise TgError(rc, "tg_material_from_angles")
        self._check(self.lib.tg_set_materials(self.ctx, member, len(problem.materials), mats))

    def _check(self, rc, nocontext=False):
        if rc != _lib.TG_OK:
            msg = "" if nocontext else self.lib.tg_last_error(self.ctx).decode()
            raise TgError(rc, msg)

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.tg_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state ------------------------------------------------------------------
    def set_params(self, phi_int_deg, phi_bed_deg):
        p = self.problem
        self.params = params_from_angles(phi_int_deg, phi_bed_deg, p.epsilon, p.gravity_scale)

    def set_state(self, h, hu, hv):
        h, hu, hv = (np.ascontiguousarray(a, dtype=np.float64) for a in (h, hu, hv))
        self._check(self.lib.tg_set_state(self.ctx, _lib.dptr(h), _lib.dptr(hu), _lib.dptr(hv)))

    def get_state(self):
        n = self.problem.n
        h, hu, hv = np.empty(n), np.empty(n), np.empty(n)
        self._check(self.lib.tg_get_state(self.ctx, _lib.dptr(h), _lib.dptr(hu), _lib.dptr(hv)))
        return h, hu, hv

    # -- forward ----------------------------------------------------------------
    def step(self):
        st = _lib.tg_stats()
        self._check(self.lib.tg_forward_step(self.ctx, C.byref(self.params), C.byref(st)))
        return st

    def run(self, nsteps=None):
        st = _lib.tg_stats()
        n = self.problem.nsteps if nsteps is None else nsteps
        self._check(self.lib.tg_forward_run(self.ctx, C.byref(self.params), n, C.byref(st)))
        return st

    def replay(self, dt_seq):
        dt = np.ascontiguousarray(dt_seq, dtype=np.float64)
        st = _lib.tg_stats()
        self._check(self.lib.tg_forward_replay(self.ctx, C.byref(self.params), len(dt),
                                               _lib.dptr(dt), C.byref(st)))
        return st

    def trajectory_state(self, s):
        """U^s of the last run (0 <= s <= its steps), all members."""
        n = self.lib.tg_num_cells(self.ctx)
        h, hu, hv = np.empty(n), np.empty(n), np.empty(n)
        self._check(self.lib.tg_get_trajectory_state(self.ctx, s, _lib.dptr(h), _lib.dptr(hu),
                                                     _lib.dptr(hv)))
        return h, hu, hv

    def write_solution_files(self, directory, keys, first=0, last=None, rank=0, order=None):
        """U^s of the last run for s in [first, last) as the reference's spill
        files directory/solution_<rank>_<s>; last None: every state."""
        k = np.ascontiguousarray(keys, dtype=np.uint32).reshape(-1)
        o = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
        if last is None:
            last = len(self.dt_record()) + 1
        self._check(self.lib.tg_write_solution_files(
            self.ctx, str(directory).encode(), rank, _lib.uptr(k),
            None if o is None else _lib.iptr(o), first, last))

    def dt_record(self):
        cnt = C.c_int()
        self._check(self.lib.tg_get_dt_record(self.ctx, None, 0, C.byref(cnt)))
        out = np.empty(max(cnt.value, 1))
        self._check(self.lib.tg_get_dt_record(self.ctx, _lib.dptr(out), cnt.value, C.byref(cnt)))
        return out[: cnt.value]

    # -- adjoint ----------------------------------------------------------------
    def _functional(self, w_flat, step):
        """tg_functional over the flat weights and per-member steps (None:
        the final step); returns it with the arrays it points into."""
        w = np.ascontiguousarray(w_flat, dtype=np.float64).reshape(-1)
        fn = _lib.tg_functional()
        fn.terminal_weight = _lib.dptr(w)
        st = None
        if step is not None:
            st = np.ascontiguousarray(np.broadcast_to(np.asarray(step, dtype=np.int32),
                                                      (self.lib.tg_num_members(self.ctx),)))
            fn.step = st.ctypes.data_as(C.POINTER(C.c_int))
        return fn, (w, st)

    def set_functional(self, terminal_weight, step=None):
        """Keeps the weights (and their step, None: the final one) on the
        device for dual_solver(None)."""
        fn, keep = self._functional(terminal_weight, step)
        self._check(self.lib.tg_set_functional(self.ctx, C.byref(fn)))

    # -- MaxH functional ----------------------------------------------------------
    def set_maxh_region(self, mask):
        """Cells of the MaxH box (n booleans, all members concatenated), or
        None to stop recording."""
        if mask is None:
            self._check(self.lib.tg_set_maxh_region(self.ctx, None))
            return
        m = np.ascontiguousarray(np.asarray(mask).reshape(-1) != 0, dtype=np.uint8)
        self._check(self.lib.tg_set_maxh_region(self.ctx,
                                                m.ctypes.data_as(C.POINTER(C.c_ubyte))))

    def box_record(self, member=0):
        """Largest h over the box in U^{s+1}, per recorded step s."""
        cnt = C.c_int()
        self._check(self.lib.tg_get_member_box_record(self.ctx, member, None, 0, C.byref(cnt)))
        out = np.empty(max(cnt.value, 1))
        self._check(self.lib.tg_get_member_box_record(self.ctx, member, _lib.dptr(out),
                                                      cnt.value, C.byref(cnt)))
        return out[: cnt.value]

    def maxh(self):
        """(hmax, step) per member: the MaxH value and the state U^step that
        first holds it (0 when the box stayed dry)."""
        nm = self.lib.tg_num_members(self.ctx)
        hm = np.empty(nm)
        st = (C.c_int * nm)()
        self._check(self.lib.tg_get_maxh(self.ctx, _lib.dptr(hm), st))
        return hm, np.array(list(st), dtype=np.int64)

    def dual_solver(self, terminal_weight, want_dphi=False, want_lambda=True, step=None):
        """Reverse sweep over the last run: returns (lambda0 (3,n) or None,
        dJ/dphi_bed or None).  terminal_weight None: the functional given to
        set_functional.  step: the state U^step the weights apply to (None:
        the final one)."""
        fnp = None
        if terminal_weight is not None:
            fn, keep = self._functional(terminal_weight, step)
            fnp = C.byref(fn)
        lam = np.empty(3 * self.problem.n) if want_lambda else None
        dphi = C.c_double(0.0)
        self._check(self.lib.tg_adjoint_run(self.ctx, C.byref(self.params), fnp,
                                            _lib.dptr(lam),
                                            C.cast(C.byref(dphi), _lib._dp) if want_dphi
                                            else None))
        return (lam.reshape(3, -1) if want_lambda else None), (dphi.value if want_dphi else None)

    def adjoint_step(self, s, dt, lam_next):
        ln = np.ascontiguousarray(lam_next, dtype=np.float64).reshape(-1)
        out = np.empty_like(ln)
        self._check(self.lib.tg_adjoint_step(self.ctx, C.byref(self.params), s, dt,
                                             _lib.dptr(ln), _lib.dptr(out)))
        return out.reshape(3, -1)

    def adjoint_step_grad(self, s, dt, lam_next):
        """One reverse step and its dJ/dphi_bed contribution."""
        ln = np.ascontiguousarray(lam_next, dtype=np.float64).reshape(-1)
        out = np.empty_like(ln)
        g = C.c_double(0.0)
        self._check(self.lib.tg_adjoint_step_grad(self.ctx, C.byref(self.params), s, dt,
                                                  _lib.dptr(ln), _lib.dptr(out),
                                                  C.cast(C.byref(g), _lib._dp)))
        return out.reshape(3, -1), g.value

    def set_error_grid(self, grid):
        """Hands an errorgrid.ErrorGrid of this context's mesh to the device
        (tg_set_error_grid)."""
        self._err_grid = grid
        self._err_struct = grid.c_struct()
        self._check(self.lib.tg_set_error_grid(self.ctx, C.byref(self._err_struct)))

    def error_indicator(self, prev, state, lam, s, dt, want_recon=False):
        """The error indicator of step s (U^s = prev -> U^{s+1} = state, dt)
        weighted with the adjoint lam, each (3, cells) in this context's
        order (tg_error_indicator): dict of residual (3, n), error (n,),
        correction (n,) per error element and, with want_recon, recon
        (9, n): reconstructed previous state, state and adjoint."""
        n = self._err_grid.n
        src = [np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
               for x in (prev, state, lam)]
        res, err, corr = np.empty(3 * n), np.empty(n), np.empty(n)
        rec = np.empty(9 * n) if want_recon else None
        self._check(self.lib.tg_error_indicator(self.ctx, C.byref(self.params), int(s), float(dt),
                                                *[_lib.dptr(x) for x in src], _lib.dptr(res),
                                                _lib.dptr(err), _lib.dptr(corr),
                                                _lib.dptr(rec)))
        out = dict(residual=res.reshape(3, n), error=err, correction=corr)
        if want_recon:
            out["recon"] = rec.reshape(9, n)
        return out

    def error_kernel_ms(self):
        """Kernel times (ms) of the last error_indicator: reconstruction,
        kact, residual."""
        ms = np.empty(3)
        self._check(self.lib.tg_error_last_ms(self.ctx, _lib.dptr(ms)))
        return ms

    def instrument(self, on=True):
        """Per-launch kernel timing and full-path cell counts (no graph
        replay while on); turning it on resets the totals."""
        self._check(self.lib.tg_set_instrumentation(self.ctx, int(on)))

    def instrumentation(self):
        """(kernel_ms[3], launches[3], full_cells[3], short_cells[3]) since
        instrument(True); kernels 0 coef, 1 update, 2 adjoint; -1 where a
        figure is not counted."""
        ms = (C.c_double * 3)()
        ln = (C.c_int64 * 3)()
        fc = (C.c_int64 * 3)()
        sc = (C.c_int64 * 3)()
        self._check(self.lib.tg_get_instrumentation(self.ctx, C.cast(ms, _lib._dp), ln, fc, sc))
        return list(ms), list(ln), list(fc), list(sc)

    def time_kernel(self, which, reps=20):
        """Average launch time (ms) of one kernel: 0 coef, 1 update, 2 adjoint."""
        ms = C.c_double()
        self._check(self.lib.tg_time_kernel(self.ctx, C.byref(self.params), which, reps,
                                            C.cast(C.byref(ms), _lib._dp)))
        return ms.value

    def _pp(self):
        return C.byref(self.params)

    def autotune(self, reps=10):
        """Times every tile variant per kernel and keeps the fastest.
        Returns (variants [coef, update, adjoint], ms (3, variants))."""
        var = (C.c_int * 3)()
        ms = np.zeros(3 * _lib.TG_NUM_TILE_VARIANTS)
        self._check(self.lib.tg_autotune(self.ctx, self._pp(), reps, var, _lib.dptr(ms)))
        return list(var), ms.reshape(3, -1)

    def tile_selection(self):
        var = (C.c_int * 3)()
        self._check(self.lib.tg_get_tile_selection(self.ctx, var))
        return list(var)

    def select_tiles(self, variants):
        v = (C.c_int * 3)(*variants)
        self._check(self.lib.tg_set_tile_selection(self.ctx, v))

    def kernel_timing(self):
        return self.timing()

    def timing(self):
        f, a, r = C.c_double(), C.c_double(), C.c_double()
        self._check(self.lib.tg_last_timing(self.ctx, C.cast(C.byref(f), _lib._dp),
                                            C.cast(C.byref(a), _lib._dp),
                                            C.cast(C.byref(r), _lib._dp)))
        return f.value, a.value, r.value

    def stored_sweep_timing(self):
        """(kernel ms, steps) of the last reverse sweep's stored-trajectory
        part: reverse steps back to back, no recomputation."""
        ms, n = C.c_double(), C.c_int()
        self._check(self.lib.tg_last_stored_sweep(self.ctx, C.cast(C.byref(ms), _lib._dp),
                                                  C.byref(n)))
        return ms.value, n.value


def _mesh_of(problem, keep):
    m = _lib.tg_mesh()
    m.kind = _lib.TG_MESH_UNIFORM
    m.nx, m.ny, m.dx, m.dy, m.x0, m.y0 = (problem.nx, problem.ny, problem.dx, problem.dy,
                                          problem.x0, problem.y0)
    g = np.ascontiguousarray(problem.grav.reshape(-1))
    dg = np.ascontiguousarray(problem.dgrav.reshape(-1))
    cv = np.ascontiguousarray(problem.curv.reshape(-1))
    keep += [g, dg, cv]
    m.grav, m.dgrav, m.curv = _lib.dptr(g), _lib.dptr(dg), _lib.dptr(cv)
    _mesh_materials(m, problem, keep)
    return m


def _mesh_materials(m, problem, keep):
    """tg_mesh.mat from the problem's per-cell material indices (or NULL)."""
    if problem.mat is None:
        return
    mat = np.ascontiguousarray(problem.mat, dtype=np.uint8).reshape(-1)
    if mat.size != problem.n:
        raise ValueError("problem.mat must hold one index per cell")
    keep.append(mat)
    m.mat = mat.ctypes.data_as(C.POINTER(C.c_uint8))


class Ensemble(Simulation):
    """Several uniform blocks (members) advanced together in every launch,
    each with its own friction angles, CFL dt and time (tg_set_ensemble)."""

    def __init__(self, problems, precision=_lib.TG_FP64, checkpoint_every=50, max_steps=None,
                 device=0, **cfg_overrides):
        self.lib = _lib.load()
        self.problems = list(problems)
        self.problem = self.problems[0]
        ms = max_steps or max(p.nsteps for p in self.problems)
        self.cfg = default_config(precision=precision, checkpoint_every=checkpoint_every,
                                  max_steps=ms, device=device, **cfg_overrides)
        self.ctx = C.c_void_p()
        self._check(self.lib.tg_create(C.byref(self.cfg), C.byref(self.ctx)), nocontext=True)
        self._arrays = []
        nm = len(self.problems)
        meshes = (_lib.tg_mesh * nm)(*[_mesh_of(p, self._arrays) for p in self.problems])
        prms = (_lib.tg_params * nm)(*[params_from_angles(p.phi_int_deg, p.phi_bed_deg,
                                                          p.epsilon, p.gravity_scale)
                                       for p in self.problems])
        self._check(self.lib.tg_set_ensemble(self.ctx, nm, meshes, prms))
        for b, p in enumerate(self.problems):
            self._set_materials(b, p)
        self.params = None
        self.offsets = []
        for b in range(nm):
            off, cnt = C.c_int64(), C.c_int64()
            self._check(self.lib.tg_member_cells(self.ctx, b, C.byref(off), C.byref(cnt)))
            self.offsets.append((off.value, cnt.value))
        self.ncell = int(self.lib.tg_num_cells(self.ctx))
        h = np.concatenate([p.h for p in self.problems])
        hu = np.concatenate([p.hu for p in self.problems])
        hv = np.concatenate([p.hv for p in self.problems])
        self.set_state(h, hu, hv)

    def _pp(self):
        return None if self.params is None else C.byref(self.params)

    def get_state(self):
        n = self.ncell
        h, hu, hv = np.empty(n), np.empty(n), np.empty(n)
        self._check(self.lib.tg_get_state(self.ctx, _lib.dptr(h), _lib.dptr(hu), _lib.dptr(hv)))
        return h, hu, hv

    def member_state(self, b):
        o, n = self.offsets[b]
        h, hu, hv = self.get_state()
        return h[o:o + n], hu[o:o + n], hv[o:o + n]

    def run(self, nsteps=None):
        nm = len(self.problems)
        st = (_lib.tg_stats * nm)()
        n = max(p.nsteps for p in self.problems) if nsteps is None else nsteps
        self._check(self.lib.tg_forward_run(self.ctx, None, n, st))
        return list(st)

    def member_dt_record(self, b):
        cnt = C.c_int()
        self._check(self.lib.tg_get_member_dt_record(self.ctx, b, None, 0, C.byref(cnt)))
        out = np.empty(max(cnt.value, 1))
        self._check(self.lib.tg_get_member_dt_record(self.ctx, b, _lib.dptr(out), cnt.value,
                                                     C.byref(cnt)))
        return out[: cnt.value]

    def concat(self, per_member):
        """(3, n_b) arrays per member -> the concatenated (3, ncell) layout."""
        return np.concatenate([np.asarray(a, dtype=np.float64).reshape(3, -1)
                               for a in per_member], axis=1)

    def set_functional(self, weights, step=None):
        """weights: list of (3, n_b) weights, kept on the device (with their
        per-member steps, None: the final one) for dual_solver(None)."""
        super().set_functional(np.ascontiguousarray(self.concat(weights)).reshape(-1), step)

    def set_maxh_region(self, masks):
        """masks: list of per-member box masks (n_b booleans), or None."""
        super().set_maxh_region(None if masks is None else
                                np.concatenate([np.asarray(m).reshape(-1) for m in masks]))

    def dual_solver(self, weights, want_dphi=False, want_lambda=True, step=None):
        """weights: list of (3, n_b) weights, or None for the functional
        given to set_functional; step: per-member states U^step they apply
        to (None: the final one).  Returns (list of lambda0 (3, n_b) or None,
        per-member dJ/dphi_bed or None)."""
        fnp = None
        if weights is not None:
            fn, keep = self._functional(self.concat(weights), step)
            fnp = C.byref(fn)
        lam = np.empty(3 * self.ncell) if want_lambda else None
        dphi = C.c_double(0.0)
        self._check(self.lib.tg_adjoint_run(self.ctx, None, fnp, _lib.dptr(lam),
                                            C.cast(C.byref(dphi), _lib._dp) if want_dphi
                                            else None))
        out = None
        if want_lambda:
            lam = lam.reshape(3, -1)
            out = [lam[:, o:o + n] for o, n in self.offsets]
        per = None
        if want_dphi:
            per = np.empty(len(self.problems))
            self._check(self.lib.tg_get_member_dphi(self.ctx, _lib.dptr(per)))
        return out, per

    def time_kernel(self, which, reps=20):
        ms = C.c_double()
        self._check(self.lib.tg_time_kernel(self.ctx, None, which, reps,
                                            C.cast(C.byref(ms), _lib._dp)))
        return ms.value
