// C ABI of the hot path: context, mesh/state transfer, the forward time loop
// with checkpoints, and the reverse (adjoint) sweep with segment recompute.
#include "../../include/tg.h"
#include "tg_kernels.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace tg;

struct tg_ctx {
  tg_config cfg{};
  hipStream_t stream = nullptr;
  std::string err;
  bool have_mesh = false, have_state = false;

  // mesh (one instance)
  std::vector<InstDev> inst_h;
  std::vector<TileDev> tiles_h;
  InstDev* inst_d = nullptr;
  TileDev* tiles_d = nullptr;
  int ninst = 0, ntiles = 0;
  long long ncell = 0;
  double dx = 0, dy = 0;

  // device fields
  double* U[3] = {nullptr, nullptr, nullptr};
  double* Ut[3] = {nullptr, nullptr, nullptr};
  double* statics = nullptr;  // 7*n
  double* kact = nullptr;
  RedDev* red = nullptr;
  double* dt_rec = nullptr;
  double* time_d = nullptr;
  int* err_d = nullptr;
  double* part_mass = nullptr;
  double* part_hmax = nullptr;
  double* dphi_part = nullptr;
  double* lam[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  double* wterm = nullptr;  // terminal weights 3n

  // forward record
  int rec_steps = 0;        // steps in the dt record
  int run_steps = 0;        // steps of the last tg_forward_run/replay (0 = none)
  int run_start_step = 0;   // dt_rec index of the run's first step
  std::vector<double*> ckpt;
  std::vector<double*> seg;
  double time_h = 0.;

  // timing
  hipEvent_t ev[6] = {};
  double fwd_ms = 0, adj_ms = 0, rec_ms = 0;
};

namespace {

#define HIPCHK(ctx, call)                                                      \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
      return TG_ERR_HIP;                                                       \
    }                                                                          \
  } while (0)

int set_err(tg_ctx* ctx, int code, const std::string& msg) {
  ctx->err = msg;
  return code;
}

template <class T>
void dfree(T*& p) {
  if (p) hipFree(p);
  p = nullptr;
}

int dalloc(tg_ctx* ctx, void** p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    ctx->err = std::string("hipMalloc failed: ") + hipGetErrorString(e);
    return TG_ERR_NOMEM;
  }
  return TG_OK;
}

void free_fields(tg_ctx* ctx) {
  for (int k = 0; k < 3; k++) {
    dfree(ctx->U[k]);
    dfree(ctx->Ut[k]);
    dfree(ctx->lam[0][k]);
    dfree(ctx->lam[1][k]);
  }
  dfree(ctx->statics);
  dfree(ctx->kact);
  dfree(ctx->part_mass);
  dfree(ctx->part_hmax);
  dfree(ctx->dphi_part);
  dfree(ctx->wterm);
  dfree(ctx->inst_d);
  dfree(ctx->tiles_d);
  for (auto& p : ctx->ckpt) dfree(p);
  for (auto& p : ctx->seg) dfree(p);
  ctx->ckpt.clear();
  ctx->seg.clear();
}

double frict_tiny(const tg_ctx* ctx, int s) {
  return (s < ctx->cfg.frict_tiny_switch) ? ctx->cfg.frict_tiny_early : ctx->cfg.frict_tiny_late;
}

void params_to_inst(tg_ctx* ctx, const tg_params* p) {
  for (auto& I : ctx->inst_h) {
    I.sin_int = p->sin_int;
    I.tan_bed = p->tan_bed;
    I.kc1 = p->kact_c1;
    I.kroot = p->kact_root;
    I.dkroot = p->dkroot_dphi;
    I.eps = p->epsilon;
    I.gscale = p->gravity_scale;
    I.maxtime = p->maxtime;
    I.nextoutput = p->nextoutput;
    I.nextsave = p->nextsave;
  }
}

int upload_params(tg_ctx* ctx, const tg_params* p) {
  if (!p) return set_err(ctx, TG_ERR_ARG, "null params");
  params_to_inst(ctx, p);
  HIPCHK(ctx, hipMemcpyAsync(ctx->inst_d, ctx->inst_h.data(), sizeof(InstDev) * ctx->ninst,
                             hipMemcpyHostToDevice, ctx->stream));
  return TG_OK;
}

StepArgs step_args(tg_ctx* ctx, double* const in[3], double* const out[3], int s) {
  StepArgs a{};
  const long long n = ctx->ncell;
  a.inst = ctx->inst_d;
  a.tiles = ctx->tiles_d;
  a.ntiles = ctx->ntiles;
  a.U = Fields{in[0], in[1], in[2]};
  const double* st = ctx->statics;
  a.S = Statics{st, st + n, st + 2 * n, st + 3 * n, st + 4 * n, st + 5 * n, st + 6 * n};
  a.kact = ctx->kact;
  a.red = ctx->red;
  a.dt_rec = ctx->dt_rec;
  a.time = ctx->time_d;
  a.max_steps = ctx->cfg.max_steps;
  a.step = s;
  a.ft = frict_tiny(ctx, s);
  a.tiny = ctx->cfg.tiny;
  a.gamma = ctx->cfg.gamma;
  a.cfl_inner = ctx->cfg.cfl_inner;
  a.cfl_outer = ctx->cfg.cfl_outer;
  a.h_out = out[0];
  a.hu_out = out[1];
  a.hv_out = out[2];
  a.err = ctx->err_d;
  return a;
}

// One forward step U=in (step index s) -> out.  compute_dt: CFL + incrtime,
// else dt_rec[s] must hold the step's dt.
void enqueue_step(tg_ctx* ctx, double* const in[3], double* const out[3], int s,
                  bool compute_dt) {
  StepArgs a = step_args(ctx, in, out, s);
  launch_coef(a, compute_dt, ctx->stream);
  if (compute_dt) launch_dt(a, ctx->ninst, ctx->stream);
  launch_update(a, ctx->stream);
}

int check_err(tg_ctx* ctx) {
  int e = 0;
  HIPCHK(ctx, hipMemcpyAsync(&e, ctx->err_d, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (e & 1) return set_err(ctx, TG_ERR_NONFINITE, "non-finite state (residual.C:137-147)");
  if (e & 2) return set_err(ctx, TG_ERR_EIGEN, "eigenvalue > 1e9 (get_coef_and_eigen.C:144)");
  return TG_OK;
}

int fill_stats(tg_ctx* ctx, tg_stats* st) {
  if (!st) return TG_OK;
  launch_stats(ctx->inst_d, ctx->tiles_d, ctx->ntiles, Fields{ctx->U[0], ctx->U[1], ctx->U[2]},
               ctx->part_mass, ctx->part_hmax, ctx->cfg.tiny, ctx->stream);
  std::vector<double> pm(ctx->ntiles), ph(ctx->ntiles);
  HIPCHK(ctx, hipMemcpyAsync(pm.data(), ctx->part_mass, sizeof(double) * ctx->ntiles,
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ph.data(), ctx->part_hmax, sizeof(double) * ctx->ntiles,
                             hipMemcpyDeviceToHost, ctx->stream));
  double dt = 0;
  if (ctx->rec_steps > 0)
    HIPCHK(ctx, hipMemcpyAsync(&dt, ctx->dt_rec + ctx->rec_steps - 1, sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
  double t = 0;
  HIPCHK(ctx, hipMemcpyAsync(&t, ctx->time_d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  double m = 0, hm = 0;
  for (int k = 0; k < ctx->ntiles; k++) {
    m += pm[k];
    hm = (ph[k] > hm) ? ph[k] : hm;
  }
  st->dt = dt;
  st->time = t;
  st->mass = m;
  st->hmax = hm;
  st->step = ctx->rec_steps;
  return TG_OK;
}

int ensure_store(tg_ctx* ctx, std::vector<double*>& v, size_t count) {
  const size_t bytes = sizeof(double) * 3 * (size_t)ctx->ncell;
  while (v.size() < count) {
    double* p = nullptr;
    int rc = dalloc(ctx, (void**)&p, bytes);
    if (rc) return rc;
    v.push_back(p);
  }
  return TG_OK;
}

int copy_state(tg_ctx* ctx, double* dst, double* const src[3]) {
  const size_t b = sizeof(double) * (size_t)ctx->ncell;
  for (int k = 0; k < 3; k++)
    HIPCHK(ctx, hipMemcpyAsync(dst + k * ctx->ncell, src[k], b, hipMemcpyDeviceToDevice,
                               ctx->stream));
  return TG_OK;
}

void split(tg_ctx* ctx, double* base, double* out[3]) {
  for (int k = 0; k < 3; k++) out[k] = base + k * ctx->ncell;
}

// Shared body of tg_forward_run / tg_forward_replay.
int forward_loop(tg_ctx* ctx, const tg_params* prm, int nsteps, const double* dt_seq,
                 tg_stats* stats) {
  if (!ctx->have_mesh || !ctx->have_state) return set_err(ctx, TG_ERR_STATE, "mesh/state not set");
  if (nsteps < 0 || nsteps > ctx->cfg.max_steps)
    return set_err(ctx, TG_ERR_ARG, "nsteps outside [0, max_steps]");
  int rc = upload_params(ctx, prm);
  if (rc) return rc;
  const int k = ctx->cfg.checkpoint_every;
  const size_t nck = (size_t)(nsteps + k - 1) / k + 1;
  rc = ensure_store(ctx, ctx->ckpt, nck);
  if (rc) return rc;
  // the record restarts: dt_rec[0..nsteps) belongs to this run
  ctx->rec_steps = 0;
  ctx->run_steps = 0;
  ctx->run_start_step = 0;
  if (dt_seq)
    HIPCHK(ctx, hipMemcpyAsync(ctx->dt_rec, dt_seq, sizeof(double) * nsteps,
                               hipMemcpyHostToDevice, ctx->stream));
  launch_reset_red(ctx->red, ctx->ninst, ctx->stream);
  HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  for (int s = 0; s < nsteps; s++) {
    if (s % k == 0) {
      rc = copy_state(ctx, ctx->ckpt[s / k], ctx->U);
      if (rc) return rc;
    }
    enqueue_step(ctx, ctx->U, ctx->Ut, s, dt_seq == nullptr);
    for (int c = 0; c < 3; c++) std::swap(ctx->U[c], ctx->Ut[c]);
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  HIPCHK(ctx, hipGetLastError());
  rc = check_err(ctx);
  if (rc) return rc;
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  ctx->fwd_ms = ms;
  ctx->rec_steps = nsteps;
  ctx->run_steps = nsteps;
  if (dt_seq) {
    double t = ctx->time_h;
    for (int s = 0; s < nsteps; s++) t += dt_seq[s];
    ctx->time_h = t;
    HIPCHK(ctx, hipMemcpyAsync(ctx->time_d, &ctx->time_h, sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
  }
  return fill_stats(ctx, stats);
}

}  // namespace

extern "C" {

void tg_config_default(tg_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->precision = TG_FP64;
  cfg->checkpoint_every = 50;
  cfg->max_steps = 4096;
  cfg->device = 0;
  cfg->tiny = 1.e-04;
  cfg->gamma = 1.0;
  cfg->cfl_inner = 0.5;
  cfg->cfl_outer = 0.5;
  cfg->frict_tiny_early = 0.1;
  cfg->frict_tiny_late = 0.000000001;
  cfg->frict_tiny_switch = 50;
  cfg->tile_x = 0;
  cfg->tile_y = 0;
}

int tg_params_from_angles(double phi_int, double phi_bed, double epsilon, double gravity_scale,
                          tg_params* out) {
  if (!out) return TG_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  // getcoef.f:39-43 evaluated once: cosphi = dcos(int), tandel = dtan(bed)
  const double cosphi = std::cos(phi_int), tandel = std::tan(phi_bed);
  out->sin_int = std::sin(phi_int);
  out->tan_bed = tandel;
  out->kact_c1 = 2.0 / (cosphi * cosphi);
  const double w = 1.0 - (1.0 + tandel * tandel) * (cosphi * cosphi);
  out->kact_root = std::sqrt(std::fabs(w));
  // d root / d phi_bed = -sign(w) cos^2 tan (1+tan^2) / root
  if (out->kact_root > 0)
    out->dkroot_dphi = -(w < 0 ? -1.0 : 1.0) * (cosphi * cosphi) * tandel *
                       (1.0 + tandel * tandel) / out->kact_root;
  out->epsilon = epsilon;
  out->gravity_scale = gravity_scale;
  out->maxtime = HUGE_VAL;
  out->nextoutput = HUGE_VAL;
  out->nextsave = HUGE_VAL;
  return TG_OK;
}

int tg_create(const tg_config* cfg, tg_ctx** out) {
  if (!out) return TG_ERR_ARG;
  *out = nullptr;
  tg_ctx* ctx = new tg_ctx();
  if (cfg) ctx->cfg = *cfg;
  else tg_config_default(&ctx->cfg);
  if (ctx->cfg.checkpoint_every < 1 || ctx->cfg.max_steps < 1) {
    delete ctx;
    return TG_ERR_ARG;
  }
  if (ctx->cfg.precision != TG_FP64) {
    delete ctx;
    return TG_ERR_UNSUPPORTED;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= ctx->cfg.device) {
    delete ctx;
    return TG_ERR_NODEVICE;
  }
  if (hipSetDevice(ctx->cfg.device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return TG_ERR_HIP;
  }
  for (auto& e : ctx->ev) hipEventCreate(&e);
  if (dalloc(ctx, (void**)&ctx->err_d, sizeof(int)) ||
      dalloc(ctx, (void**)&ctx->dt_rec, sizeof(double) * ctx->cfg.max_steps)) {
    tg_destroy(ctx);
    return TG_ERR_NOMEM;
  }
  hipMemset(ctx->err_d, 0, sizeof(int));
  *out = ctx;
  return TG_OK;
}

void tg_destroy(tg_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->cfg.device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  free_fields(ctx);
  dfree(ctx->red);
  dfree(ctx->time_d);
  dfree(ctx->err_d);
  dfree(ctx->dt_rec);
  for (auto& e : ctx->ev)
    if (e) hipEventDestroy(e);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* tg_last_error(const tg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int tg_set_mesh(tg_ctx* ctx, const tg_mesh* m) {
  if (!ctx || !m) return TG_ERR_ARG;
  if (m->kind != TG_MESH_UNIFORM)
    return set_err(ctx, TG_ERR_UNSUPPORTED, "only TG_MESH_UNIFORM is built");
  if (m->nx < 1 || m->ny < 1 || !(m->dx > 0) || !(m->dy > 0) || !m->grav || !m->dgrav || !m->curv)
    return set_err(ctx, TG_ERR_ARG, "bad mesh");
  // The kernels use dx, dy for the centre spacing; the reference takes
  // differences of centre coordinates (element2.C:894-895).  Both agree only
  // if those differences are exact, which holds for dyadic spacings and
  // origins; reject meshes where they do not.
  for (int i = 0; i + 1 < m->nx; i++)
    if ((m->x0 + (i + 1) * m->dx) - (m->x0 + i * m->dx) != m->dx)
      return set_err(ctx, TG_ERR_ARG, "x centre spacing is not exactly dx");
  for (int j = 0; j + 1 < m->ny; j++)
    if ((m->y0 + (j + 1) * m->dy) - (m->y0 + j * m->dy) != m->dy)
      return set_err(ctx, TG_ERR_ARG, "y centre spacing is not exactly dy");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  free_fields(ctx);
  ctx->have_mesh = ctx->have_state = false;
  ctx->rec_steps = ctx->run_steps = 0;
  const long long n = (long long)m->nx * m->ny;
  ctx->ncell = n;
  ctx->ninst = 1;
  ctx->dx = m->dx;
  ctx->dy = m->dy;
  InstDev I{};
  I.nx = m->nx;
  I.ny = m->ny;
  I.off = 0;
  I.dx = m->dx;
  I.dy = m->dy;
  ctx->inst_h.assign(1, I);
  ctx->tiles_h.clear();
  for (int ty = 0; ty < m->ny; ty += TILE_Y)
    for (int tx = 0; tx < m->nx; tx += TILE_X) ctx->tiles_h.push_back(TileDev{0, tx, ty, 0});
  ctx->ntiles = (int)ctx->tiles_h.size();
  int rc = 0;
  const size_t nb = sizeof(double) * n;
  for (int k = 0; k < 3 && !rc; k++) {
    rc = dalloc(ctx, (void**)&ctx->U[k], nb);
    if (!rc) rc = dalloc(ctx, (void**)&ctx->Ut[k], nb);
    if (!rc) rc = dalloc(ctx, (void**)&ctx->lam[0][k], nb);
    if (!rc) rc = dalloc(ctx, (void**)&ctx->lam[1][k], nb);
  }
  if (!rc) rc = dalloc(ctx, (void**)&ctx->statics, 7 * nb);
  if (!rc) rc = dalloc(ctx, (void**)&ctx->kact, nb);
  if (!rc) rc = dalloc(ctx, (void**)&ctx->wterm, 3 * nb);
  if (!rc) rc = dalloc(ctx, (void**)&ctx->part_mass, sizeof(double) * ctx->ntiles);
  if (!rc) rc = dalloc(ctx, (void**)&ctx->part_hmax, sizeof(double) * ctx->ntiles);
  if (!rc) rc = dalloc(ctx, (void**)&ctx->dphi_part, sizeof(double) * ctx->ntiles);
  if (!rc) rc = dalloc(ctx, (void**)&ctx->inst_d, sizeof(InstDev));
  if (!rc) rc = dalloc(ctx, (void**)&ctx->tiles_d, sizeof(TileDev) * ctx->ntiles);
  if (!rc && !ctx->red) rc = dalloc(ctx, (void**)&ctx->red, sizeof(RedDev) * 1);
  if (!rc && !ctx->time_d) rc = dalloc(ctx, (void**)&ctx->time_d, sizeof(double) * 1);
  if (rc) {
    free_fields(ctx);
    return rc;
  }
  HIPCHK(ctx, hipMemcpyAsync(ctx->statics, m->grav, 3 * nb, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->statics + 3 * n, m->dgrav, 2 * nb, hipMemcpyHostToDevice,
                             ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->statics + 5 * n, m->curv, 2 * nb, hipMemcpyHostToDevice,
                             ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->tiles_d, ctx->tiles_h.data(), sizeof(TileDev) * ctx->ntiles,
                             hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->inst_d, ctx->inst_h.data(), sizeof(InstDev),
                             hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->have_mesh = true;
  return TG_OK;
}

int tg_set_state(tg_ctx* ctx, const double* h, const double* hu, const double* hv) {
  if (!ctx || !h || !hu || !hv) return TG_ERR_ARG;
  if (!ctx->have_mesh) return set_err(ctx, TG_ERR_STATE, "no mesh");
  const size_t nb = sizeof(double) * ctx->ncell;
  HIPCHK(ctx, hipMemcpyAsync(ctx->U[0], h, nb, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->U[1], hu, nb, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->U[2], hv, nb, hipMemcpyHostToDevice, ctx->stream));
  ctx->time_h = 0.;
  HIPCHK(ctx, hipMemcpyAsync(ctx->time_d, &ctx->time_h, sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(ctx->err_d, 0, sizeof(int), ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->rec_steps = ctx->run_steps = 0;
  ctx->have_state = true;
  return TG_OK;
}

int tg_get_state(tg_ctx* ctx, double* h, double* hu, double* hv) {
  if (!ctx || !h || !hu || !hv) return TG_ERR_ARG;
  if (!ctx->have_state) return set_err(ctx, TG_ERR_STATE, "no state");
  const size_t nb = sizeof(double) * ctx->ncell;
  HIPCHK(ctx, hipMemcpyAsync(h, ctx->U[0], nb, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(hu, ctx->U[1], nb, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(hv, ctx->U[2], nb, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return TG_OK;
}

int tg_forward_step(tg_ctx* ctx, const tg_params* prm, tg_stats* stats) {
  if (!ctx) return TG_ERR_ARG;
  if (!ctx->have_mesh || !ctx->have_state) return set_err(ctx, TG_ERR_STATE, "mesh/state not set");
  if (ctx->rec_steps >= ctx->cfg.max_steps) return set_err(ctx, TG_ERR_ARG, "dt record full");
  int rc = upload_params(ctx, prm);
  if (rc) return rc;
  if (ctx->rec_steps == 0) launch_reset_red(ctx->red, ctx->ninst, ctx->stream);
  enqueue_step(ctx, ctx->U, ctx->Ut, ctx->rec_steps, true);
  for (int c = 0; c < 3; c++) std::swap(ctx->U[c], ctx->Ut[c]);
  HIPCHK(ctx, hipGetLastError());
  rc = check_err(ctx);
  if (rc) return rc;
  ctx->rec_steps++;
  ctx->run_steps = 0;  // a single step does not keep checkpoints
  HIPCHK(ctx, hipMemcpyAsync(&ctx->time_h, ctx->time_d, sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return fill_stats(ctx, stats);
}

int tg_forward_run(tg_ctx* ctx, const tg_params* prm, int nsteps, tg_stats* stats) {
  if (!ctx) return TG_ERR_ARG;
  int rc = forward_loop(ctx, prm, nsteps, nullptr, stats);
  if (!rc) {
    HIPCHK(ctx, hipMemcpyAsync(&ctx->time_h, ctx->time_d, sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  }
  return rc;
}

int tg_forward_replay(tg_ctx* ctx, const tg_params* prm, int nsteps, const double* dt_seq,
                      tg_stats* stats) {
  if (!ctx || (!dt_seq && nsteps > 0)) return TG_ERR_ARG;
  return forward_loop(ctx, prm, nsteps, dt_seq, stats);
}

int tg_get_dt_record(tg_ctx* ctx, double* dt_out, int capacity, int* count) {
  if (!ctx || !count) return TG_ERR_ARG;
  *count = ctx->rec_steps;
  if (dt_out && capacity > 0) {
    int m = capacity < ctx->rec_steps ? capacity : ctx->rec_steps;
    HIPCHK(ctx, hipMemcpyAsync(dt_out, ctx->dt_rec, sizeof(double) * m, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  }
  return TG_OK;
}

static AdjArgs adj_args(tg_ctx* ctx, double* const U[3], double* const lnext[3],
                        double* const lout[3], int s, bool dphi) {
  AdjArgs a{};
  const long long n = ctx->ncell;
  a.inst = ctx->inst_d;
  a.tiles = ctx->tiles_d;
  a.ntiles = ctx->ntiles;
  a.U = Fields{U[0], U[1], U[2]};
  const double* st = ctx->statics;
  a.S = Statics{st, st + n, st + 2 * n, st + 3 * n, st + 4 * n, st + 5 * n, st + 6 * n};
  a.lam_next = Fields{lnext[0], lnext[1], lnext[2]};
  a.dt_rec = ctx->dt_rec;
  a.max_steps = ctx->cfg.max_steps;
  a.step = s;
  a.ft = frict_tiny(ctx, s);
  a.tiny = ctx->cfg.tiny;
  a.gamma = ctx->cfg.gamma;
  a.lam_h = lout[0];
  a.lam_hu = lout[1];
  a.lam_hv = lout[2];
  a.dphi_part = dphi ? ctx->dphi_part : nullptr;
  return a;
}

int tg_adjoint_run(tg_ctx* ctx, const tg_params* prm, const tg_functional* fn,
                   double* lambda0_out, double* dJ_dphibed_out) {
  if (!ctx || !fn || !fn->terminal_weight) return TG_ERR_ARG;
  if (ctx->run_steps <= 0) return set_err(ctx, TG_ERR_STATE, "no forward run recorded");
  int rc = upload_params(ctx, prm);
  if (rc) return rc;
  const long long n = ctx->ncell;
  const int N = ctx->run_steps, k = ctx->cfg.checkpoint_every;
  rc = ensure_store(ctx, ctx->seg, (size_t)k);
  if (rc) return rc;
  const bool dphi = dJ_dphibed_out != nullptr;
  if (dphi) HIPCHK(ctx, hipMemsetAsync(ctx->dphi_part, 0, sizeof(double) * ctx->ntiles, ctx->stream));
  // lambda^N = func_sens at the final step (calc_adjoint.C:93-97)
  const size_t nb = sizeof(double) * n;
  for (int c = 0; c < 3; c++)
    HIPCHK(ctx, hipMemcpyAsync(ctx->lam[0][c], fn->terminal_weight + c * n, nb,
                               hipMemcpyHostToDevice, ctx->stream));
  int cur = 0;
  float rec_ms = 0, ms = 0;
  HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
  const int nseg = (N + k - 1) / k;
  for (int sg = nseg - 1; sg >= 0; sg--) {
    const int s0 = sg * k, s1 = (s0 + k < N) ? s0 + k : N;
    // recompute U^{s0..s1-1} from the checkpoint with the recorded dt
    HIPCHK(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
    double* base[3];
    split(ctx, ctx->ckpt[sg], base);
    rc = copy_state(ctx, ctx->seg[0], base);
    if (rc) return rc;
    for (int s = s0; s + 1 < s1; s++) {
      double* in[3];
      double* out[3];
      split(ctx, ctx->seg[s - s0], in);
      split(ctx, ctx->seg[s - s0 + 1], out);
      enqueue_step(ctx, in, out, s, false);
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
    // reverse steps s = s1-1 .. s0
    for (int s = s1 - 1; s >= s0; s--) {
      double* Us[3];
      split(ctx, ctx->seg[s - s0], Us);
      AdjArgs a = adj_args(ctx, Us, ctx->lam[cur], ctx->lam[1 - cur], s, dphi);
      launch_adjoint(a, ctx->stream);
      cur = 1 - cur;
    }
    HIPCHK(ctx, hipEventSynchronize(ctx->ev[5]));
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]));
    rec_ms += ms;
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
  HIPCHK(ctx, hipGetLastError());
  rc = check_err(ctx);
  if (rc) return rc;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
  ctx->adj_ms = ms;
  ctx->rec_ms = rec_ms;
  if (lambda0_out)
    for (int c = 0; c < 3; c++)
      HIPCHK(ctx, hipMemcpyAsync(lambda0_out + c * n, ctx->lam[cur][c], nb,
                                 hipMemcpyDeviceToHost, ctx->stream));
  if (dphi) {
    std::vector<double> part(ctx->ntiles);
    HIPCHK(ctx, hipMemcpyAsync(part.data(), ctx->dphi_part, sizeof(double) * ctx->ntiles,
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    double acc = 0;
    for (double v : part) acc += v;
    *dJ_dphibed_out = acc;
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return TG_OK;
}

int tg_adjoint_step(tg_ctx* ctx, const tg_params* prm, int s, double dt,
                    const double* lambda_next, double* lambda_out) {
  if (!ctx || !lambda_next || !lambda_out || s < 0 || s >= ctx->cfg.max_steps)
    return TG_ERR_ARG;
  if (!ctx->have_state) return set_err(ctx, TG_ERR_STATE, "no state");
  int rc = upload_params(ctx, prm);
  if (rc) return rc;
  const long long n = ctx->ncell;
  const size_t nb = sizeof(double) * n;
  // the step's dt goes into the record slot the kernel reads
  HIPCHK(ctx, hipMemcpyAsync(ctx->dt_rec + s, &dt, sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
  for (int c = 0; c < 3; c++)
    HIPCHK(ctx, hipMemcpyAsync(ctx->lam[0][c], lambda_next + c * n, nb, hipMemcpyHostToDevice,
                               ctx->stream));
  AdjArgs a = adj_args(ctx, ctx->U, ctx->lam[0], ctx->lam[1], s, false);
  launch_adjoint(a, ctx->stream);
  HIPCHK(ctx, hipGetLastError());
  for (int c = 0; c < 3; c++)
    HIPCHK(ctx, hipMemcpyAsync(lambda_out + c * n, ctx->lam[1][c], nb, hipMemcpyDeviceToHost,
                               ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->run_steps = 0;  // the dt record slot was overwritten
  return TG_OK;
}

int tg_last_timing(tg_ctx* ctx, double* forward_ms, double* adjoint_ms, double* recompute_ms) {
  if (!ctx) return TG_ERR_ARG;
  if (forward_ms) *forward_ms = ctx->fwd_ms;
  if (adjoint_ms) *adjoint_ms = ctx->adj_ms;
  if (recompute_ms) *recompute_ms = ctx->rec_ms;
  return TG_OK;
}

int tg_time_kernel(tg_ctx* ctx, const tg_params* prm, int which, int reps, double* avg_ms) {
  if (!ctx || !avg_ms || reps < 1 || which < 0 || which > 2) return TG_ERR_ARG;
  if (!ctx->have_mesh || !ctx->have_state || ctx->rec_steps < 1)
    return set_err(ctx, TG_ERR_STATE, "needs a completed forward run");
  int rc = upload_params(ctx, prm);
  if (rc) return rc;
  // every variant reads U and writes scratch (Ut or lam[1]); step 0's dt
  StepArgs a = step_args(ctx, ctx->U, ctx->Ut, 0);
  AdjArgs b = adj_args(ctx, ctx->U, ctx->lam[0], ctx->lam[1], 0, false);
  launch_coef(a, false, ctx->stream);  // kact of the current state
  HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  for (int r = 0; r < reps; r++) {
    if (which == 0) launch_coef(a, false, ctx->stream);
    else if (which == 1) launch_update(a, ctx->stream);
    else launch_adjoint(b, ctx->stream);
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventSynchronize(ctx->ev[1]));
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  *avg_ms = ms / reps;
  return check_err(ctx);
}

int64_t tg_num_cells(const tg_ctx* ctx) { return ctx ? (int64_t)ctx->ncell : 0; }

}  // extern "C"
