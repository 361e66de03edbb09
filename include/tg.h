/*
 * tg.h -- C ABI of the MI355X granular-flow hot path (forward finite-volume
 * step + discrete-adjoint reverse sweep).  Plain C types only.
 *
 * Every entry point replaces a pass body of the reference (cited per call).
 * Conventions (the reference has no status codes; these replace its
 * printf+exit/assert error handling, residual.C:137-147):
 *   - every call returns TG_OK (0) or a negative tg_status;
 *   - one host thread per context; calls are ordered on the context's
 *     stream; getters synchronise;
 *   - host-side arrays are FP64 regardless of the internal precision;
 *   - transcendentals of the material parameters are evaluated on the host
 *     (tg_params_from_angles) so the device never re-derives them.
 */
#ifndef TG_H
#define TG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tg_status {
  TG_OK = 0,
  TG_ERR_ARG = -1,          /* bad argument / size                           */
  TG_ERR_HIP = -2,          /* HIP runtime error                             */
  TG_ERR_STATE = -3,        /* call out of order (no mesh, no forward run)   */
  TG_ERR_NOMEM = -4,        /* device allocation failed                      */
  TG_ERR_NONFINITE = -5,    /* NaN/Inf in the state (residual.C:137-147)     */
  TG_ERR_EIGEN = -6,        /* eigenvalue > 1e9 (get_coef_and_eigen.C:144)   */
  TG_ERR_UNSUPPORTED = -7,  /* mesh kind / mode not built                    */
  TG_ERR_NODEVICE = -8,     /* no HIP device                                 */
  TG_ERR_TOO_LARGE = -9     /* block larger than tg_step_noalloc's fixed
                               scratch (TG_NOALLOC_MAX_CELLS)                */
} tg_status;

typedef enum tg_precision {
  TG_FP64 = 0,   /* every value FP64: the accuracy reference                  */
  TG_MIXED = 1,  /* FP32 fluxes, slopes, sources and Jacobians (increments);
                    FP32 static coefficients and kact; state and adjoint kept
                    in FP64 and advanced by the FP32 increments; FP64 thin-
                    layer velocity recovery, CFL reduction, mass and
                    gradient sums                                           */
  TG_FP32 = 2,   /* as TG_MIXED but the state and adjoint are stored in FP32
                    (half the bytes; larger error over long runs)           */
  TG_FWD64_ADJ32 = 3  /* forward sweep and its reverse recompute exactly as
                    TG_FP64 (bit-identical states, dt and functional); the
                    adjoint's Jacobian and gradient arithmetic in FP32 with
                    every branch decision taken in FP64 on the FP64 state,
                    adjoint stored and accumulated in FP64, FP32 copy of the
                    static coefficients for the adjoint                     */
} tg_precision;

/* Context configuration.  tg_config_default() fills the reference's values. */
typedef struct tg_config {
  int precision;          /* tg_precision                                      */
  int checkpoint_every;   /* k: forward states stored every k steps (>= 1)     */
  int max_steps;          /* capacity of the dt[] record                      */
  int device;             /* HIP device ordinal                                */
  double tiny;            /* GEOFLOW_TINY = 1e-4 (constant.h:68)               */
  double gamma;           /* slope limiter gamma = 1 (hpfem.C:82)              */
  double cfl_inner;       /* 0.5 (get_coef_and_eigen.C:168)                    */
  double cfl_outer;       /* 0.5 (get_coef_and_eigen.C:175)                    */
  double frict_tiny_early;/* 0.1  (hpfem.C:232-235)                            */
  double frict_tiny_late; /* 1e-9 (hpfem.C:232-235)                            */
  int frict_tiny_switch;  /* steps 0..49 use the early value (hpfem.C:232,
                             calc_jacobain.C:81-84)                            */
  int tile_x, tile_y;     /* 0 x 0: built-in tile variant (tg_autotune may
                             re-select); 64x4, 32x8 or 16x16: that shape,
                             XCD-banded, for every kernel                     */
  int use_graph;          /* 1: tg_forward_run / tg_forward_replay and tg_adjoint_run
                             replay a hipGraph of the whole sweep (captured on
                             the first call with a given step count, tile
                             selection and buffer parity); 0: launch kernel
                             by kernel.  Results are identical.              */
  int64_t store_bytes;    /* device memory for keeping the last states of a
                             forward run whole (0: checkpoints only).  The
                             reverse sweep reads those states and recomputes
                             only below them.  Results are identical.        */
} tg_config;

/* Mesh.  TG_MESH_UNIFORM: an nx*ny block of equal cells; cell c = j*nx + i,
 * centre (x0 + i*dx, y0 + j*dy); cells on the outer ring have the domain
 * boundary (neigh_proc == INIT) on their outer sides.  Static per-cell
 * coefficients (element2.C:3545-3703) are FP64 SoA arrays of n cells.
 *
 * TG_MESH_LEAVES: the n leaves of a 2:1-balanced adaptive mesh, in the
 * caller's order (e.g. the element-table traversal of tg_host_flatten), as
 * the reference's Element neighbour fields (element2.h:652-667): sides
 * 0 south, 1 east, 2 north, 3 west (positive_x_side = 1); slot s is the
 * neighbour on side s (with two finer neighbours, the one next to corner s)
 * and slot s+4 the second neighbour (next to corner s+1) or a copy of slot
 * s.  nproc is neigh_proc of one process: -1 domain boundary (nbr -1), -2 no
 * second neighbour, 0 present.  A leaf may not have the domain boundary on
 * two opposite sides.  nx, ny, dx, dy, x0, y0 are ignored; the leaf path is
 * built for TG_FP64 (other precisions return TG_ERR_UNSUPPORTED) and for a
 * single member. */
typedef enum tg_mesh_kind { TG_MESH_UNIFORM = 0, TG_MESH_LEAVES = 1 } tg_mesh_kind;

typedef struct tg_mesh {
  int kind;
  int nx, ny;
  double dx, dy;
  double x0, y0;
  const double* grav;   /* [3*n]: gravity x | y | z  (calc_gravity_vector)   */
  const double* dgrav;  /* [2*n]: d gravity_z / dx | dy (calc_d_gravity)     */
  const double* curv;   /* [2*n]: curvature x | y (calc_topo_data)           */
  /* TG_MESH_LEAVES only */
  int n;                /* number of leaves                                  */
  const int32_t* nbr;   /* [8*n] Element::neighbor as leaf indices           */
  const int32_t* nproc; /* [8*n] Element::neigh_proc                         */
  const int32_t* gen;   /* [n]   Element::generation                         */
  const double* cdx;    /* [2*n] Element::dx (cell size x, y)                */
  const double* xc;     /* [2*n] Element::coord (centre x, y)                */
  /* Both kinds: material index of every cell / leaf (Element::material,
   * read by get_coef_and_eigen.C:112-114 and corrector.C:108 as
   * bedfrict[material]), 0 .. TG_MAX_MATERIALS-1; NULL = one material.
   * The reference numbers its materials from 1 (GIS raster ids,
   * element2.C:3690): pass id - 1. */
  const uint8_t* mat;
} tg_mesh;

/* Material and time parameters of one run (MatProps, TimeProps). */
typedef struct tg_params {
  double sin_int;        /* sin(phi_int)           residual.C:208            */
  double tan_bed;        /* tan(phi_bed)           residual.C:213            */
  double kact_c1;        /* 2/cos^2(phi_int)       getcoef.f:40              */
  double kact_root;      /* sqrt|1-(1+tan^2 phi_bed) cos^2 phi_int| getcoef.f:41 */
  double dkroot_dphi;    /* d kact_root / d phi_bed (bed-friction gradient)  */
  double epsilon;        /* HEIGHT_SCALE / LENGTH_SCALE                      */
  double gravity_scale;  /* GRAVITY_SCALE                                    */
  double maxtime;        /* non-dimensional end time (incrtime clamp)        */
  double nextoutput;     /* next output time (incrtime clamp); +inf = none   */
  double nextsave;       /* next save time (incrtime clamp); +inf = none     */
} tg_params;

/* Bed friction of one material (MatProps::bedfrict[material] with the
 * internal friction of tg_params): tan(phi_bed) (residual.C:213),
 * sqrt|1 - (1 + tan^2 phi_bed) cos^2 phi_int| (getcoef.f:41) and its
 * phi_bed derivative (bed-friction gradient). */
#define TG_MAX_MATERIALS 16
typedef struct tg_material {
  double tan_bed;
  double kact_root;
  double dkroot_dphi;
} tg_material;

/* Mass/height statistics of the last completed step (step.C:205 realvolume,
 * get_coef_and_eigen.C:102-104 hmax); FP64 sums on the device. */
typedef struct tg_stats {
  double dt;       /* dt taken by the step                                   */
  double time;     /* simulated non-dimensional time after the step          */
  double mass;     /* sum dx*dy*h over all cells after the step              */
  double hmax;     /* max h over wet cells after the step                    */
  int step;        /* number of steps completed                              */
} tg_stats;

/* Functional J = sum_cells w . U^{step_b} of each member b (weights FP64 SoA
 * [3*n]): the reference's per-step func_sens vector, nonzero at one step
 * (calc_adjoint.C:49, 93-97, 150; MaxH_Functional::clac_func_sens
 * calc_adjoint.C:470-496 sets it at the MaxH step only).  step: per member,
 * in [0, steps of the last forward run]; NULL = the final step of the run. */
typedef struct tg_functional {
  const double* terminal_weight;
  const int* step;
} tg_functional;

typedef struct tg_ctx tg_ctx;

/* ---- setup --------------------------------------------------------------- */

/* Fills cfg with the reference's constants (constant.h, hpfem.C). */
void tg_config_default(tg_config* cfg);

/* MatProps setup: computes the host-side transcendentals once
 * (MatProps ctor properties.h:578-608; getcoef.f:39-43).  Angles in radians. */
int tg_params_from_angles(double phi_int, double phi_bed, double epsilon,
                          double gravity_scale, tg_params* out);

/* One material's host-side transcendentals (frict.data read, datread.C:379-403,
 * with getcoef.f:39-43).  Angles in radians. */
int tg_material_from_angles(double phi_int, double phi_bed, tg_material* out);

/* Creates a context on cfg->device (owns device buffers and stream). */
int tg_create(const tg_config* cfg, tg_ctx** out);
void tg_destroy(tg_ctx* ctx);
const char* tg_last_error(const tg_ctx* ctx);

/* Copies mesh tables and static coefficients to the device (replaces
 * setup_geoflow's per-element data, setup_geoflow.C:24; called per re-mesh). */
int tg_set_mesh(tg_ctx* ctx, const tg_mesh* mesh);

/* Materials 0 .. nmat-1 of a member whose mesh gave per-cell materials
 * (tg_mesh.mat): the MatProps::bedfrict table (datread.C:379-403).  A
 * material index no call gave takes the member's own tg_params values (the
 * ones tg_set_ensemble or the run's prm hold).  dJ/dphi_bed of
 * tg_adjoint_run is then the derivative with respect to one common change of
 * every material's bed friction angle (the sum of the per-material
 * derivatives).  TG_ERR_STATE if the mesh has no material indices. */
int tg_set_materials(tg_ctx* ctx, int member, int nmat, const tg_material* mats);

/* State U = (h, hu, hv), FP64 host arrays of n (Element::state_vars).
 * With several members, the arrays hold all cells: member b's cells at
 * [offset_b, offset_b + count_b) (tg_member_cells). */
int tg_set_state(tg_ctx* ctx, const double* h, const double* hu, const double* hv);
int tg_get_state(tg_ctx* ctx, double* h, double* hu, double* hv);

/* The simulation clock of the state just set (tg_set_state resets it to step
 * 0, time 0): the state is U^{step0} at simulated time `time`, so a run's
 * step s is the simulation's step step0 + s (frict_tiny schedule,
 * hpfem.C:232-235) and incrtime clamps against time + elapsed
 * (properties.h:397-410).  A time loop that re-meshes (hpfem.C:244-258)
 * hands each mesh's segment over with its clock. */
int tg_set_clock(tg_ctx* ctx, int step0, double time);

/* ---- ensembles ----------------------------------------------------------
 * Several independent uniform blocks (members, e.g. samples of a stochastic
 * study, lhsbed.C / lhsvol.C) advanced together: every kernel launch covers
 * all members' cells, each member with its own parameters, CFL dt and time.
 * The reference runs each sample as a separate MPI job (lhs.readme:6-14);
 * this has no counterpart there.  After tg_set_ensemble, the prm argument
 * of tg_forward_run / tg_adjoint_run may be NULL (keep per-member params);
 * a non-NULL prm overrides every member.  Functional weights, lambda0 and
 * states use the concatenated layout of tg_set_state. */
int tg_set_ensemble(tg_ctx* ctx, int nmembers, const tg_mesh* meshes, const tg_params* params);
int tg_num_members(const tg_ctx* ctx);
int tg_member_cells(const tg_ctx* ctx, int member, int64_t* offset, int64_t* count);
int tg_set_member_state(tg_ctx* ctx, int member, const double* h, const double* hu,
                        const double* hv);
int tg_get_member_state(tg_ctx* ctx, int member, double* h, double* hu, double* hv);
int tg_get_member_dt_record(tg_ctx* ctx, int member, double* dt_out, int capacity, int* count);
/* Per-member dJ/d phi_bed of the last tg_adjoint_run that asked for it. */
int tg_get_member_dphi(tg_ctx* ctx, double* dJ_dphibed);

/* ---- forward ---------------------------------------------------------------- */

/* One order-1 step: replaces step() step.C:29-275 (slopes, get_coef_and_eigen,
 * incrtime, calc_edge_states, corrector).  Appends dt to the dt[] record. */
int tg_forward_step(tg_ctx* ctx, const tg_params* prm, tg_stats* stats);

/* nsteps steps (the hpfem.C:222-300 time loop without adaptation): stores a
 * checkpoint of U every cfg->checkpoint_every steps and the dt[] record for
 * the adjoint sweep.  Restarts the record (U at entry is checkpoint 0). */
int tg_forward_run(tg_ctx* ctx, const tg_params* prm, int nsteps, tg_stats* stats);

/* Same time loop with a prescribed dt sequence (frozen-dt replay, used for
 * finite-difference checks of the adjoint; calc_jacobain.C:89 uses the
 * recorded dt the same way). */
int tg_forward_replay(tg_ctx* ctx, const tg_params* prm, int nsteps,
                      const double* dt_seq, tg_stats* stats);

/* U^s of the last forward run, 0 <= s <= its steps (the states
 * SolRec::record_solution keeps, dualmesh.C:35-78): read from the stored
 * trajectory, or recomputed from the checkpoint below s with the recorded
 * dt.  FP64 host arrays of n. */
int tg_get_trajectory_state(tg_ctx* ctx, int s, double* h, double* hu, double* hv);

/* Writes U^s of the last forward run for s in [first, last) as the
 * reference's spill files <dir>/solution_<rank %04d>_<s %08d>
 * (SolRec::wrtie_sol_to_disk, dualmesh.C:87-142; layout as
 * tg_host_write_solution): keys [2n] the element keys, order as there
 * (NULL = array order).  last may be the run's steps + 1 (U^N included). */
int tg_write_solution_files(tg_ctx* ctx, const char* dir, int rank, const uint32_t* keys,
                            const int32_t* order, int first, int last);

/* Copies the dt[] record of the last run (TimeProps::dt). */
int tg_get_dt_record(tg_ctx* ctx, double* dt_out, int capacity, int* count);

/* ---- adjoint ---------------------------------------------------------------- */

/* Reverse sweep over the last forward run: replaces dual_solver.C:183-257
 * with calc_jacobian + calc_adjoint at every step (not only the MaxH step).
 * Forward states are recomputed segment by segment from the checkpoints.
 * lambda0_out: FP64 [3*n] adjoint at step 0 (= dJ/dU^0), may be NULL.
 * dJ_dphibed_out: dJ/d phi_bed at frozen dt (may be NULL).
 * fn NULL: the functional last given to tg_set_functional (weights and steps).
 * Members whose step lies below the largest step start from lambda = 0 and
 * take their weights when the sweep reaches their step. */
int tg_adjoint_run(tg_ctx* ctx, const tg_params* prm, const tg_functional* fn,
                   double* lambda0_out, double* dJ_dphibed_out);

/* Keeps a copy of the functional's terminal weights on the device, for
 * tg_adjoint_run with fn NULL (no host upload per sweep).  No reference
 * counterpart (the reference evaluates func_sens on its own mesh). */
int tg_set_functional(tg_ctx* ctx, const tg_functional* fn);

/* MaxH functional (MaxH_Functional, calc_adjoint.C:420-548, adjoint.h:270-297):
 * mask [n] marks the functional's box (nonzero = cell centre strictly inside,
 * calc_adjoint.C:462-464); NULL switches the record off.  While set, every
 * step of tg_forward_run / tg_forward_replay / tg_forward_step records the
 * largest h > 0 over each member's box cells in the state it produced
 * (check_maxh, calc_adjoint.C:443-465); a run restarts the record. */
int tg_set_maxh_region(tg_ctx* ctx, const unsigned char* mask);
/* The record of member b: hmax_out[s] = largest h in the box of U^{s+1}
 * (0 when none is > 0), s < count (the steps recorded); the per-step values
 * MaxH_Functional::check_maxh compares (calc_adjoint.C:443-465). */
int tg_get_member_box_record(tg_ctx* ctx, int member, double* hmax_out, int capacity,
                             int* count);
/* Per member: the largest recorded h and the first step at which the box
 * holds it (check_maxh's strict '>' over the steps in order, from 0;
 * MaxH_Functional::update calc_adjoint.C:498-514 for one process): the
 * state U^{step}, 0 when the box never held h > 0.  With weights (1, 0, 0)
 * on the box cells and this step in tg_functional, tg_adjoint_run is the
 * reverse sweep of the MaxH functional. */
int tg_get_maxh(tg_ctx* ctx, double* hmax, int* step);

/* One reverse step on the current state: given U^s (the context state) and
 * lambda^{s+1}, writes lambda^s (FP64 SoA [3*n]); step index s selects
 * frict_tiny; dt is the step's dt.  Mirrors calc_jacobian + calc_adjoint of
 * one dual_solver iteration. */
int tg_adjoint_step(tg_ctx* ctx, const tg_params* prm, int s, double dt,
                    const double* lambda_next, double* lambda_out);

/* As tg_adjoint_step, and also returns this step's contribution
 * lambda^{s+1} . dU^{s+1}/d phi_bed to dJ/d phi_bed, per member
 * (dJ_dphibed [number of members]); the terms tg_adjoint_run sums. */
int tg_adjoint_step_grad(tg_ctx* ctx, const tg_params* prm, int s, double dt,
                         const double* lambda_next, double* lambda_out, double* dJ_dphibed);

/* ---- error indicator on the error grid ----------------------------------------
 * The reference's Error build (dual_solver.C:143-175, 208-240) estimates
 * the error of the forward solution on the error grid -- the mesh with
 * every leaf refined once (uinform_refine, uniform_refine.C:24-130) -- from
 * bilinear reconstructions of the state, the previous state and the
 * adjoint (bilinear_interp, bilinear_interp.C:126-556): the residual of one
 * step of the reconstructed previous state against the reconstructed state
 * (update_bilinear_error_grid, error_grid_routines.C:13-53, and
 * error_compute, error_compute.C:24-100), weighted by the adjoint.
 *
 * tg_set_error_grid hands over the error grid of the context's mesh (one
 * member): the error elements' leaf tables as for tg_mesh (nbr, nproc, gen,
 * cdx, xc; host layout), their static coefficients (grav, dgrav, curv; mat
 * with the materials of the context's material table, NULL = one
 * material), the mesh cell each one refines (father) and the interpolation
 * stencils of tg_host_error_stencil (tg_host.h: kind, pts, pxy).  The grid
 * lives until the next tg_set_mesh / tg_set_error_grid. */
typedef struct tg_error_grid {
  int n;                  /* error elements                                    */
  const int32_t* nbr;     /* [8*n] as tg_mesh                                  */
  const int32_t* nproc;   /* [8*n]                                             */
  const int32_t* gen;     /* [n]                                               */
  const double* cdx;      /* [2*n]                                             */
  const double* xc;       /* [2*n]                                             */
  const double* grav;     /* [3*n]                                             */
  const double* dgrav;    /* [2*n]                                             */
  const double* curv;     /* [2*n]                                             */
  const uint8_t* mat;     /* [n] or NULL                                       */
  const int32_t* father;  /* [n] cell of the context's mesh                    */
  const int32_t* kind;    /* [n] TG_ERR_INTERP_* rule                          */
  const int32_t* pts;     /* [4*n] mesh cells of the stencil points            */
  const double* pxy;      /* [8*n] their coordinates                           */
} tg_error_grid;
int tg_set_error_grid(tg_ctx* ctx, const tg_error_grid* g);

/* One evaluation for the step U^s -> U^{s+1} of the context's mesh: prev =
 * U^s, state = U^{s+1}, lambda = the adjoint the estimate is weighted with,
 * each FP64 SoA [3 * cells] (h | hu | hv) on the host; step s selects
 * frict_tiny as tg_forward_step does (error_compute.C:63-66), dt the step's
 * dt; prm the material constants.  Outputs (host, any may be NULL), per
 * error element: residual [3*n] (predicted - reconstructed state,
 * ErrorElem::residual), error [n] (sum_j |res_j (lambda_j - lambda~_j)|,
 * el_error[1]), correction [n] (sum_j res_j lambda_j), recon [9*n] (the
 * reconstructed previous state, state and adjoint, bilin_prev_state,
 * bilin_state, bilin_adj).  lambda_j is the element's father's adjoint
 * (send_from_dual_to_error, error_grid_routines.C:130-170), lambda~ its
 * reconstruction.  Synchronous.  Returns TG_OK, TG_ERR_ARG, TG_ERR_STATE
 * (no error grid) or TG_ERR_HIP. */
int tg_error_indicator(tg_ctx* ctx, const tg_params* prm, int step, double dt,
                       const double* prev, const double* state, const double* lambda,
                       double* residual, double* error, double* correction, double* recon);
/* Kernel times of the last tg_error_indicator in milliseconds (HIP events):
 * reconstruction, kact, residual. */
int tg_error_last_ms(const tg_ctx* ctx, double ms[3]);

/* ---- one step without allocation ----------------------------------------------
 * No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  tg_step_noalloc is
 * tg_forward_step (one order-1 step, step.C:29-275) on caller-owned device
 * arrays with every intermediate -- kact, the CFL reduction, dt, the tile
 * tables and the update's tile list -- in a fixed device array of the
 * library (TG_NOALLOC_MAX_CELLS cells at most, about 9 MiB per device, held
 * until the process ends).  It allocates and frees nothing.  Results equal
 * tg_forward_step's on the same state bit for bit.  Calls on one device share
 * the array: calls that overlap on one device race (the call synchronises the
 * stream before it returns).
 *   blk:     nx x ny uniform block, cell spacing dx, dy; statics = device
 *            array [7 * nx * ny] FP64: gravity x | y | z, d gravity_z dx | dy,
 *            curvature x | y (tg_mesh grav, dgrav, curv in that order);
 *   cfg:     thresholds, CFL factors, frict_tiny schedule and device
 *            (tg_config; other fields ignored);
 *   step:    step index s (selects frict_tiny); time: simulated time before
 *            the step (incrtime clamp against prm's maxtime / output / save);
 *   d_h ...: device arrays [nx * ny] of U^s; d_h_out ...: U^{s+1} (distinct
 *            arrays); stream: a hipStream_t or NULL (the null stream);
 *   dt_out:  the dt taken (host).
 * Returns TG_OK, TG_ERR_ARG, TG_ERR_TOO_LARGE (more than
 * TG_NOALLOC_MAX_CELLS cells), TG_ERR_NONFINITE, TG_ERR_EIGEN or TG_ERR_HIP. */
#define TG_NOALLOC_MAX_CELLS (1 << 20)
typedef struct tg_noalloc_block {
  int nx, ny;
  double dx, dy;
  const double* statics;
} tg_noalloc_block;
int tg_step_noalloc(const tg_config* cfg, const tg_noalloc_block* blk, const tg_params* prm,
                    int step, double time, const double* d_h, const double* d_hu,
                    const double* d_hv, double* d_h_out, double* d_hu_out, double* d_hv_out,
                    void* stream, double* dt_out);
/* Bytes of the fixed device array tg_step_noalloc uses. */
long long tg_noalloc_scratch_bytes(void);

/* ---- timing (HIP events around the context's kernels) ---------------------- */

/* Time of the last forward run and reverse sweep in milliseconds, from HIP
 * events on the context's stream around their launches (graph replays or
 * kernel by kernel): forward_ms the forward sweep, adjoint_ms the whole
 * reverse sweep, recompute_ms its checkpoint-segment recomputation (restore
 * and forward steps), so adjoint_ms - recompute_ms is the reverse steps. */
int tg_last_timing(tg_ctx* ctx, double* forward_ms, double* adjoint_ms,
                   double* recompute_ms);
/* The part of the last reverse sweep that read the stored trajectory (no
 * recomputation: reverse steps back to back, and the clearing of the
 * gradient sums): its kernel time in milliseconds and its number of steps.
 * No reference counterpart (measurement). */
int tg_last_stored_sweep(tg_ctx* ctx, double* stored_ms, int* stored_steps);

/* Average duration of one launch of a single kernel, timed with HIP events
 * on the context's stream over `reps` back-to-back launches on the current
 * state (no reference counterpart: measurement only).  which: 0 = forward
 * coefficient/CFL pass, 1 = forward update pass, 2 = adjoint gather (with
 * its bed-friction gradient term and, in the FP32-adjoint modes, its
 * thin-layer pass).
 * Requires a completed tg_forward_run (and tg_adjoint_run for which = 2). */
typedef enum tg_kernel_id { TG_K_COEF = 0, TG_K_UPDATE = 1, TG_K_ADJOINT = 2 } tg_kernel_id;
int tg_time_kernel(tg_ctx* ctx, const tg_params* prm, int which, int reps, double* avg_ms);

/* Instrumentation (no reference counterpart: measurement only).  While on,
 * forward runs and reverse sweeps launch kernel by kernel (no graph replay),
 * every coefficient / update / adjoint launch is timed with a HIP event pair,
 * and the kernels count the cells that take their full path and the cells
 * they write on a short path.  Uniform blocks: the coefficient pass's full
 * path is a cell above tiny, its short path a cell of a visited tile at or
 * below tiny (reads h); the update's and adjoint's full path is a cell of a
 * tile that is not entirely dry (update: h < tiny on the tile and its side
 * neighbours, state copied) or void (adjoint: h == 0 and lambda^{s+1} == 0
 * on the tile, lambda^s = 0 written), which are the short path.  Cells of
 * TG_FP64 tiles known to be +0.0 from the zero marks are neither (nothing
 * read or written).  Leaves: full path = h above tiny (coefficients), not
 * dry (update), h != 0 (adjoint).  Turning it on resets the totals. */
int tg_set_instrumentation(tg_ctx* ctx, int on);
/* Totals since instrumentation was turned on, per kernel id: summed launch
 * time (ms), launches, full-path cells and short-path cells (-1 for figures
 * a kernel does not count: the mixed-precision forward kernels count
 * neither, the mixed-precision adjoint and the leaf kernels no short-path
 * cells). */
int tg_get_instrumentation(tg_ctx* ctx, double* kernel_ms, int64_t* launches,
                           int64_t* full_cells, int64_t* short_cells);

/* ---- tile shapes (no reference counterpart: launch configuration) ----------
 * Every kernel runs one 256-thread workgroup per tile of cells.  Built tile
 * variants v = 0..TG_NUM_TILE_VARIANTS-1: a tile shape (tile_x x tile_y
 * cells) and a tile order -- row-major, or banded by XCD (workgroup b runs
 * on XCD b % 8, and a banded table gives each XCD a contiguous run of tile
 * rows so halo rows stay in that XCD's L2).  Results do not depend on the
 * variant except for the order of FP64 partial sums (mass statistics,
 * dJ/dphi_bed), which is fixed for a fixed variant. */
#define TG_NUM_TILE_VARIANTS 4
int tg_tile_variant(int variant, int* tile_x, int* tile_y, int* xcd_banded);

/* Times every variant for each kernel (coef, update, adjoint; `reps` launches
 * each on the current state, after a completed tg_forward_run) and selects
 * the fastest per kernel for the rest of the context's life.
 * variant_out [3] (may be NULL); ms_out [3 * TG_NUM_TILE_VARIANTS] average
 * launch times (may be NULL).  Overwrites the adjoint buffers. */
int tg_autotune(tg_ctx* ctx, const tg_params* prm, int reps, int* variant_out, double* ms_out);

/* Variants in use per kernel (coef, update, adjoint); set them explicitly
 * (e.g. to reproduce an earlier autotune). */
int tg_get_tile_selection(const tg_ctx* ctx, int* variant_out);
int tg_set_tile_selection(tg_ctx* ctx, const int* variant);

/* Free and total device memory of a HIP device (for sizing store_bytes). */
int tg_device_memory(int device, int64_t* free_bytes, int64_t* total_bytes);

/* Cells in the context's mesh (sum over instances). */
int64_t tg_num_cells(const tg_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
