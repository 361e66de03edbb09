// C ABI of the hot path: context, mesh/state transfer, the forward time loop
// with checkpoints, and the reverse (adjoint) sweep with segment recompute.
// A context holds one or several uniform blocks ("members"); every kernel
// launch covers all members (tile table), each member with its own material
// parameters, CFL dt and time.
#include "../../include/tg.h"
#include "../../include/tg_host.h"
#include "tg_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

using namespace tg;

// An instantiated capture of a whole sweep and the inputs it was captured
// for (step count, buffer parity, tile variants, ...).
struct SweepGraph {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  long long key[8] = {};
  bool valid = false;
};

struct tg_ctx {
  tg_config cfg{};
  hipStream_t stream = nullptr;
  std::string err;
  bool have_mesh = false, have_state = false;
  int mode = TG_FP64;
  int es = 8;   // bytes per stored state / adjoint value
  int ses = 8;  // bytes per static coefficient and kact value

  // members
  std::vector<InstDev> inst_h;
  // one tile table per variant (SHAPE_TX / SHAPE_XCD); sel[k] = variant
  // used by kernel k (0 coef, 1 update, 2 adjoint); ntiles = largest table
  std::vector<TileDev> tiles_h[NSHAPE];
  InstDev* inst_d = nullptr;
  TileDev* tiles_d[NSHAPE] = {};
  int nt[NSHAPE] = {};
  // mark blocks of each variant's tile table (uniform blocks)
  std::vector<BlockDev> blocks_h[NSHAPE];
  BlockDev* blocks_d[NSHAPE] = {};
  int nb[NSHAPE] = {};
  int sel[3] = {0, 0, 0};
  int ninst = 0, ntiles = 0;
  long long ncell = 0;

  // device fields (element type double or float by es).  Every state and
  // adjoint buffer is one allocation of state_bytes: h, hu, hv (es * ncell
  // bytes each, h first) and, at mk_off, the zero marks (nsuper words) that
  // the FP64 uniform kernels keep when `marks` is on.
  size_t state_bytes = 0, mk_off = 0;
  long long nsuper = 0;
  bool marks = false;
  // tile shape (SHAPE_TX) whose tiles the mark bytes of the state buffers
  // [0] and of the adjoint buffers [1] describe: the update's and the
  // adjoint's variant when they were last written (0: none yet)
  int mk_tx[2] = {0, 0};
  // the autotuner's timing launches: single-tile skips off, so every
  // variant is timed on the same footing (marks are kept in one shape only)
  bool bytes_off = false;
  void* U[3] = {nullptr, nullptr, nullptr};
  void* Ut[3] = {nullptr, nullptr, nullptr};
  void* U_home = nullptr;  // U[0] as allocated
  void* statics = nullptr;  // 7 * ncell, element size ses
  void* statics_f = nullptr;  // 7 * ncell FP32 copy for the FP32 adjoint of TG_FWD64_ADJ32
  void* kact = nullptr;
  void* lam[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  void* fw = nullptr;  // terminal weights kept on the device (tg_set_functional)
  bool have_fw = false;
  std::vector<int> fw_steps;  // their per-member steps (empty: the final step)
  void* wsrc = nullptr;  // host weights of a sweep that places them at several steps
  // MaxH functional: each member's box cells in chunks, and the per-step
  // record of the largest h over them (ninst * max_steps)
  BoxChunk* box_chunks = nullptr;
  long long* box_cells = nullptr;
  int nbox_chunks = 0;
  bool box_on = false;  // a region is set (it may hold no cell)
  long long box_version = 0;  // changes with every tg_set_maxh_region
  double* box_rec = nullptr;
  RedDev* red = nullptr;
  double* dt_rec = nullptr;  // ninst * max_steps
  double* time_d = nullptr;  // ninst
  int* err_d = nullptr;
  double* part_mass = nullptr;
  double* part_hmax = nullptr;
  // per-member reductions (k_member_reduce): each variant's tiles grouped by
  // member (member b: red_idx[v][red_off[v][b] .. red_off[v][b+1]), row-major
  // within the member), the results on the device and in pinned host memory
  int* red_off[NSHAPE] = {};
  int* red_idx[NSHAPE] = {};
  double* red_out = nullptr;   // [4 * ninst]
  double* red_host = nullptr;  // [4 * ninst], pinned
  // leaf path: the step and state buffer whose dual kact the last reverse
  // launch made ahead (-1: none) and whether with its phi_bed derivative
  int kd_step = -1;
  void* kd_U = nullptr;
  bool kd_dk = false;
  double* dphi_part = nullptr;
  int2* thin_list = nullptr;  // FP32 adjoint modes: cells gathered in FP64
  int* thin_count = nullptr;
  // tiles taking the full path of the update / adjoint (compacted by their
  // classifying pass) and the list length
  int* act_list = nullptr;
  int* act_count = nullptr;

  // adaptive mesh (TG_MESH_LEAVES): leaf tables and reverse-step buffers
  bool leaves = false;
  LeafDev leaf{};
  int* leaf_nbr = nullptr;
  uint32_t* leaf_meta = nullptr;
  int* leaf_cnb_off = nullptr;  // chunks holding neighbours of each chunk (CSR)
  int* leaf_cnb = nullptr;
  double* leaf_cdx = nullptr;
  double* leaf_xc = nullptr;
  LeafBufs lb{};

  // error grid (tg_set_error_grid): its tables, static coefficients and
  // interpolation stencils, and the buffers of one tg_error_indicator call
  int err_n = 0;
  LeafDev err_leaf{};
  int* err_nbr = nullptr;
  uint32_t* err_meta = nullptr;
  double* err_cdx = nullptr;
  double* err_xc = nullptr;
  double* err_statics = nullptr;  // 7 * err_n: gx, gy, gz, dgx, dgy, kx, ky
  uint8_t* err_mat = nullptr;
  int* err_father = nullptr;
  int* err_kind = nullptr;
  int* err_pts = nullptr;
  double* err_pxy = nullptr;
  double* err_fwd = nullptr;  // 9 * ncell: previous state, state, adjoint
  double* err_buf = nullptr;  // 15 * err_n: reconstruction 9, kact 1, residual 3, error, correction
  hipEvent_t err_ev[4] = {};
  double err_ms[3] = {-1., -1., -1.};

  // per-cell materials (tg_mesh.mat, tg_set_materials): the cells' material
  // index (null: one material per member) and the material table, MAX_MAT
  // entries per member; entries at or above mat_count[b] follow member b's
  // own parameters
  uint8_t* mat_d = nullptr;
  MatDev* mtab_d = nullptr;
  std::vector<MatDev> mtab_h;
  std::vector<int> mat_count;

  // step index of the state tg_set_state gave (tg_set_clock): the run's
  // step s is step step0 + s of the simulation (frict_tiny schedule)
  int step0 = 0;

  // forward record
  int rec_steps = 0;  // steps in the dt record
  int run_steps = 0;  // steps of the last forward run/replay (0 = none)
  std::vector<void*> ckpt, seg;
  // stored trajectory: U^{traj_s0 .. run_steps-1} of the last run, kept
  // whole (cfg.store_bytes) so the reverse sweep recomputes only below it
  void* traj_block = nullptr;  // one allocation of traj_cap states
  size_t traj_cap = 0;
  int traj_s0 = 0;
  std::vector<double> member_dphi;

  // captured sweeps (cfg.use_graph), re-captured when their key changes:
  // the forward sweep, the reverse sweep's stored part, and per checkpoint
  // segment its recompute and its adjoint steps (gr[sg], gs[sg]; launched
  // one after the other with events between them, so the recompute and the
  // adjoint steps are timed apart inside a replayed sweep)
  SweepGraph gf, ga;
  std::vector<SweepGraph> gr, gs;
  bool gf_end_is_b = false;  // the captured forward sweep leaves U^N in Ut
  std::vector<hipEvent_t> seg_ev;  // two per segment: around its recompute

  // timing
  hipEvent_t ev[7] = {};
  double fwd_ms = 0, adj_ms = 0, rec_ms = 0;
  double stored_ms = 0;  // reverse sweep over the stored trajectory only
  int stored_steps = 0;

  // instrumentation (tg_set_instrumentation): an event pair around every
  // coefficient / update / adjoint launch, full-path cell counters
  bool instrument = false;
  unsigned long long* active_d = nullptr;  // [6]: full-path, then short-path cells
  std::vector<hipEvent_t> iev;             // pool, two per recorded launch
  std::vector<int> iev_kind;               // kernel of each recorded pair
  double inst_ms[3] = {0, 0, 0};
  long long inst_launches[3] = {0, 0, 0};
  bool inst_counted[3] = {true, true, true};
  bool inst_short_counted[3] = {true, true, true};
};

namespace {

#define HIPCHK(ctx, call)                                             \
  do {                                                                \
    hipError_t e_ = (call);                                           \
    if (e_ != hipSuccess) {                                           \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_); \
      return TG_ERR_HIP;                                              \
    }                                                                 \
  } while (0)

int set_err(tg_ctx* ctx, int code, const std::string& msg) {
  ctx->err = msg;
  return code;
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);  // teardown: nothing to report to
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

void drop_graph(SweepGraph& G) {
  if (G.x) (void)hipGraphExecDestroy(G.x);  // teardown: nothing to report to
  if (G.g) (void)hipGraphDestroy(G.g);
  G = SweepGraph{};
}

void drop_segment_graphs(tg_ctx* ctx) {
  for (auto& G : ctx->gr) drop_graph(G);
  for (auto& G : ctx->gs) drop_graph(G);
  ctx->gr.clear();
  ctx->gs.clear();
}

void free_error_grid(tg_ctx* ctx) {
  dfree(ctx->err_nbr);
  dfree(ctx->err_meta);
  dfree(ctx->err_cdx);
  dfree(ctx->err_xc);
  dfree(ctx->err_statics);
  dfree(ctx->err_mat);
  dfree(ctx->err_father);
  dfree(ctx->err_kind);
  dfree(ctx->err_pts);
  dfree(ctx->err_pxy);
  dfree(ctx->err_fwd);
  dfree(ctx->err_buf);
  ctx->err_n = 0;
  ctx->err_leaf = LeafDev{};
}

void free_fields(tg_ctx* ctx) {
  drop_graph(ctx->gf);
  drop_graph(ctx->ga);
  drop_segment_graphs(ctx);
  dfree(ctx->U[0]);  // one allocation per buffer, h first
  dfree(ctx->Ut[0]);
  dfree(ctx->lam[0][0]);
  dfree(ctx->lam[1][0]);
  for (int k = 0; k < 3; k++) ctx->U[k] = ctx->Ut[k] = ctx->lam[0][k] = ctx->lam[1][k] = nullptr;
  dfree(ctx->fw);
  ctx->have_fw = false;
  ctx->fw_steps.clear();
  dfree(ctx->wsrc);
  dfree(ctx->box_chunks);
  dfree(ctx->box_cells);
  dfree(ctx->box_rec);
  ctx->nbox_chunks = 0;
  ctx->box_on = false;
  dfree(ctx->statics);
  dfree(ctx->statics_f);
  dfree(ctx->kact);
  dfree(ctx->part_mass);
  dfree(ctx->mat_d);
  dfree(ctx->mtab_d);
  dfree(ctx->part_hmax);
  dfree(ctx->red_out);
  for (auto& t : ctx->red_off) dfree(t);
  for (auto& t : ctx->red_idx) dfree(t);
  if (ctx->red_host) (void)hipHostFree(ctx->red_host);
  ctx->red_host = nullptr;
  dfree(ctx->dphi_part);
  dfree(ctx->thin_list);
  dfree(ctx->thin_count);
  dfree(ctx->act_list);
  dfree(ctx->act_count);
  free_error_grid(ctx);
  dfree(ctx->leaf_nbr);
  dfree(ctx->leaf_meta);
  dfree(ctx->leaf_cnb_off);
  dfree(ctx->leaf_cnb);
  dfree(ctx->lb.kdmk);
  dfree(ctx->leaf_cdx);
  dfree(ctx->leaf_xc);
  dfree(ctx->lb.kd);
  dfree(ctx->lb.dkd);
  dfree(ctx->lb.jac);
  ctx->leaf = LeafDev{};
  ctx->leaves = false;
  dfree(ctx->inst_d);
  for (auto& t : ctx->tiles_d) dfree(t);
  for (auto& t : ctx->blocks_d) dfree(t);
  dfree(ctx->red);
  dfree(ctx->time_d);
  dfree(ctx->dt_rec);
  for (auto& p : ctx->ckpt) dfree(p);
  for (auto& p : ctx->seg) dfree(p);
  dfree(ctx->traj_block);
  ctx->traj_cap = 0;
  ctx->ckpt.clear();
  ctx->seg.clear();
}

double frict_tiny(const tg_ctx* ctx, int s) {
  return (s + ctx->step0 < ctx->cfg.frict_tiny_switch) ? ctx->cfg.frict_tiny_early
                                                       : ctx->cfg.frict_tiny_late;
}

// the first step index of a run at which frict_tiny takes its late value
// (what captured sweeps depend on besides their length)
long long frict_switch_local(const tg_ctx* ctx) {
  const long long v = (long long)ctx->cfg.frict_tiny_switch - ctx->step0;
  return v < 0 ? 0 : v;
}

void params_into(InstDev& I, const tg_params* p) {
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

void marks_follow_selection(tg_ctx* ctx);

// prm == nullptr keeps each member's own parameters (tg_set_ensemble)
int upload_params(tg_ctx* ctx, const tg_params* p) {
  if (ctx->marks) marks_follow_selection(ctx);
  if (p)
    for (auto& I : ctx->inst_h) params_into(I, p);
  HIPCHK(ctx, hipMemcpyAsync(ctx->inst_d, ctx->inst_h.data(), sizeof(InstDev) * ctx->ninst,
                             hipMemcpyHostToDevice, ctx->stream));
  if (ctx->mat_d) {  // materials no tg_set_materials gave follow the member's parameters
    for (int b = 0; b < ctx->ninst; b++) {
      const InstDev& I = ctx->inst_h[b];
      for (int m = ctx->mat_count[b]; m < MAX_MAT; m++)
        ctx->mtab_h[(size_t)b * MAX_MAT + m] = MatDev{I.tan_bed, I.kroot, I.dkroot, 0.};
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->mtab_d, ctx->mtab_h.data(),
                               sizeof(MatDev) * ctx->mtab_h.size(), hipMemcpyHostToDevice,
                               ctx->stream));
  }
  return TG_OK;
}

template <class T>
T* at(void* base, long long k, long long n) {
  return (T*)base + k * n;
}

// FP64 forward (bit-identical to TG_FP64) in these modes
bool fwd64(const tg_ctx* ctx) { return ctx->mode == TG_FP64 || ctx->mode == TG_FWD64_ADJ32; }

Statics32 statics32(tg_ctx* ctx) {
  const long long n = ctx->ncell;
  void* b = ctx->statics_f ? ctx->statics_f : ctx->statics;
  return Statics32{at<float>(b, 0, n), at<float>(b, 1, n), at<float>(b, 2, n),
                   at<float>(b, 3, n), at<float>(b, 4, n), at<float>(b, 5, n),
                   at<float>(b, 6, n)};
}

// points a kernel's arguments at the tile table of its selected variant
template <class A>
void use_tiles(const tg_ctx* ctx, int kernel, A& a) {
  a.mat = ctx->mat_d;
  a.mtab = ctx->mtab_d;
  const int v = ctx->sel[kernel];
  a.tiles = ctx->tiles_d[v];
  a.ntiles = ctx->nt[v];
  a.blocks = ctx->blocks_d[v];
  a.nblocks = ctx->nb[v];
  a.shape = v;
}

// Zero marks of a buffer given by its field pointers (h first), or null
// when the kernels keep none.
uint32_t* marks_of(const tg_ctx* ctx, void* const f[3]) {
  return ctx->marks ? (uint32_t*)((char*)f[0] + ctx->mk_off) : nullptr;
}
// Marks of a buffer just written other than by the marking kernels:
// nothing known.
void clear_marks(tg_ctx* ctx, void* base) {
  if (ctx->nsuper > 0)
    launch_zero((char*)base + ctx->mk_off, sizeof(uint32_t) * (size_t)ctx->nsuper, ctx->stream);
}

// Every state and adjoint buffer's marks: nothing known.
void clear_all_marks(tg_ctx* ctx) {
  void* bufs[] = {ctx->U[0], ctx->Ut[0], ctx->lam[0][0], ctx->lam[1][0], ctx->fw, ctx->wsrc};
  for (void* b : bufs)
    if (b) clear_marks(ctx, b);
  for (void* b : ctx->ckpt)
    if (b) clear_marks(ctx, b);
  for (void* b : ctx->seg)
    if (b) clear_marks(ctx, b);
  for (size_t i = 0; i < ctx->traj_cap; i++)
    clear_marks(ctx, (char*)ctx->traj_block + i * ctx->state_bytes);
}

// A mark byte describes one tile of the writer's shape.  State buffers are
// written by the update pass and adjoint buffers by the adjoint pass; when
// the selected shapes change, bytes written in the old shapes are dropped
// (whole-word marks mean the same in every shape, single bytes do not).
void marks_follow_selection(tg_ctx* ctx) {
  const int tx[2] = {SHAPE_TX[ctx->sel[1]], SHAPE_TX[ctx->sel[2]]};
  if (tx[0] == ctx->mk_tx[0] && tx[1] == ctx->mk_tx[1]) return;
  clear_all_marks(ctx);
  ctx->mk_tx[0] = tx[0];
  ctx->mk_tx[1] = tx[1];
}

// Instrumentation: resolves the recorded event pairs into the per-kernel
// totals and frees the pool slots.
int inst_flush(tg_ctx* ctx) {
  if (ctx->iev_kind.empty()) return TG_OK;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (size_t i = 0; i < ctx->iev_kind.size(); i++) {
    float ms = 0;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->iev[2 * i], ctx->iev[2 * i + 1]));
    ctx->inst_ms[ctx->iev_kind[i]] += ms;
    ctx->inst_launches[ctx->iev_kind[i]] += 1;
  }
  ctx->iev_kind.clear();
  return TG_OK;
}

// Event recorded before (begin) or after a launch of kernel `kind`; false
// when instrumentation is off.  A failed pool refill turns it off.
void inst_mark(tg_ctx* ctx, int kind, bool begin) {
  if (!ctx->instrument) return;
  if (begin) {
    const size_t need = 2 * (ctx->iev_kind.size() + 1);
    if (need > ctx->iev.size()) {
      if (ctx->iev.size() >= 8192 && inst_flush(ctx) == TG_OK) {
        // pool reused after the flush
      } else {
        for (int k = 0; k < 256; k++) {
          hipEvent_t e = nullptr;
          if (hipEventCreate(&e) != hipSuccess) { ctx->instrument = false; return; }
          ctx->iev.push_back(e);
        }
      }
    }
    ctx->iev_kind.push_back(kind);
    (void)hipEventRecord(ctx->iev[2 * (ctx->iev_kind.size() - 1)], ctx->stream);
  } else {
    (void)hipEventRecord(ctx->iev[2 * (ctx->iev_kind.size() - 1) + 1], ctx->stream);
  }
}

unsigned long long* inst_counter(tg_ctx* ctx, int kind, bool counts) {
  if (!ctx->instrument) return nullptr;
  if (!counts) ctx->inst_counted[kind] = false;
  return counts ? ctx->active_d + kind : nullptr;
}
// counter of the cells a launch writes on a short path (null when not
// instrumenting; kernels that do not count leave the figure unknown)
unsigned long long* inst_short(tg_ctx* ctx, int kind, bool counts) {
  if (!ctx->instrument) return nullptr;
  if (!counts) ctx->inst_short_counted[kind] = false;
  return counts ? ctx->active_d + 3 + kind : nullptr;
}

template <class S>
StepArgsM<S> step_args_m(tg_ctx* ctx, void* const in[3], void* const out[3], int s) {
  StepArgsM<S> a{};
  a.inst = ctx->inst_d;
  a.U = FieldsT<S>{(S*)in[0], (S*)in[1], (S*)in[2]};
  a.S32 = statics32(ctx);
  a.kact = (float*)ctx->kact; a.red = ctx->red; a.dt_rec = ctx->dt_rec; a.time = ctx->time_d;
  a.max_steps = ctx->cfg.max_steps; a.step = s; a.ft = frict_tiny(ctx, s);
  a.tiny = ctx->cfg.tiny; a.gamma = ctx->cfg.gamma;
  a.h_out = (S*)out[0]; a.hu_out = (S*)out[1]; a.hv_out = (S*)out[2];
  a.err = ctx->err_d;
  a.mk_in = marks_of(ctx, in);
  a.mk_out = marks_of(ctx, out);
  return a;
}

template <class S>
void enqueue_step_m(tg_ctx* ctx, void* const in[3], void* const out[3], int s, bool compute_dt) {
  StepArgsM<S> a = step_args_m<S>(ctx, in, out, s);
  use_tiles(ctx, 0, a);
  inst_counter(ctx, 0, false);
  inst_counter(ctx, 1, false);
  inst_short(ctx, 0, false);
  inst_short(ctx, 1, false);
  inst_mark(ctx, 0, true);
  a.zbytes = !ctx->bytes_off && SHAPE_TX[ctx->sel[0]] == ctx->mk_tx[0];
  launch_coef_m<S>(a, compute_dt, ctx->stream);
  inst_mark(ctx, 0, false);
  if (compute_dt) {
    StepArgs d{};
    d.inst = ctx->inst_d; d.red = ctx->red; d.dt_rec = ctx->dt_rec; d.time = ctx->time_d;
    d.max_steps = ctx->cfg.max_steps; d.step = s;
    d.cfl_inner = ctx->cfg.cfl_inner; d.cfl_outer = ctx->cfg.cfl_outer;
    launch_dt(d, ctx->ninst, ctx->stream);
  }
  use_tiles(ctx, 1, a);
  inst_mark(ctx, 1, true);
  a.zbytes = !ctx->bytes_off && SHAPE_TX[ctx->sel[1]] == ctx->mk_tx[0];
  launch_update_m<S>(a, ctx->stream);
  inst_mark(ctx, 1, false);
}

template <class S>
void enqueue_adjoint_m(tg_ctx* ctx, void* const U[3], void* const lnext[3], void* const lout[3],
                       int s, bool dphi) {
  AdjArgsM<S> a{};
  a.inst = ctx->inst_d;
  use_tiles(ctx, 2, a);
  a.U = FieldsT<S>{(S*)U[0], (S*)U[1], (S*)U[2]};
  a.S32 = statics32(ctx);
  a.lam_next = FieldsT<S>{(S*)lnext[0], (S*)lnext[1], (S*)lnext[2]};
  a.dt_rec = ctx->dt_rec; a.max_steps = ctx->cfg.max_steps; a.step = s;
  a.ft = frict_tiny(ctx, s); a.tiny = ctx->cfg.tiny;
  a.gamma = ctx->cfg.gamma;
  a.lam_h = (S*)lout[0]; a.lam_hu = (S*)lout[1]; a.lam_hv = (S*)lout[2];
  a.dphi_part = dphi ? ctx->dphi_part : nullptr;
  a.thin_list = ctx->thin_list;
  a.thin_count = ctx->thin_count;
  a.s64 = (ctx->ses == 8) ? (const double*)ctx->statics : nullptr;
  a.ncell = ctx->ncell;
  a.tile_list = ctx->act_list;
  a.tile_count = ctx->act_count;
  a.mk_U = marks_of(ctx, U);
  a.mk_lnext = marks_of(ctx, lnext);
  a.mk_out = marks_of(ctx, lout);
  a.u_bytes = !ctx->bytes_off && SHAPE_TX[ctx->sel[2]] == ctx->mk_tx[0];
  a.l_bytes = !ctx->bytes_off && SHAPE_TX[ctx->sel[2]] == ctx->mk_tx[1];
  a.active = inst_counter(ctx, 2, true);
  inst_short(ctx, 2, false);
  inst_mark(ctx, 2, true);
  launch_adjoint_m<S>(a, ctx->stream);
  inst_mark(ctx, 2, false);
}

void enqueue_step(tg_ctx* ctx, void* const in[3], void* const out[3], int s, bool compute_dt) {
  const long long n = ctx->ncell;
  if (fwd64(ctx)) {
    StepArgs a{};
    a.inst = ctx->inst_d;
    a.U = Fields{(double*)in[0], (double*)in[1], (double*)in[2]};
    a.S = Statics{at<double>(ctx->statics, 0, n), at<double>(ctx->statics, 1, n),
                  at<double>(ctx->statics, 2, n), at<double>(ctx->statics, 3, n),
                  at<double>(ctx->statics, 4, n), at<double>(ctx->statics, 5, n),
                  at<double>(ctx->statics, 6, n)};
    a.kact = (double*)ctx->kact; a.red = ctx->red; a.dt_rec = ctx->dt_rec; a.time = ctx->time_d;
    a.max_steps = ctx->cfg.max_steps; a.step = s; a.ft = frict_tiny(ctx, s);
    a.tiny = ctx->cfg.tiny; a.gamma = ctx->cfg.gamma;
    a.cfl_inner = ctx->cfg.cfl_inner; a.cfl_outer = ctx->cfg.cfl_outer;
    a.h_out = (double*)out[0]; a.hu_out = (double*)out[1]; a.hv_out = (double*)out[2];
    a.err = ctx->err_d;
    a.tile_list = ctx->act_list;
    a.tile_count = ctx->act_count;
    a.mk_in = marks_of(ctx, in);
    a.mk_out = marks_of(ctx, out);
    a.zbytes = !ctx->bytes_off && SHAPE_TX[ctx->sel[1]] == ctx->mk_tx[0];
    if (ctx->leaves) {
      a.mat = ctx->mat_d;
      a.mtab = ctx->mtab_d;
      a.active = inst_counter(ctx, 0, true);
      a.short_cells = inst_short(ctx, 0, true);
      inst_mark(ctx, 0, true);
      launch_leaf_coef(a, ctx->leaf, compute_dt, ctx->stream);
      inst_mark(ctx, 0, false);
      if (compute_dt) launch_dt(a, 1, ctx->stream);
      a.active = inst_counter(ctx, 1, true);
      a.short_cells = inst_short(ctx, 1, true);
      inst_mark(ctx, 1, true);
      launch_leaf_update(a, ctx->leaf, ctx->stream);
      inst_mark(ctx, 1, false);
      return;
    }
    use_tiles(ctx, 0, a);
    a.zbytes = !ctx->bytes_off && SHAPE_TX[ctx->sel[0]] == ctx->mk_tx[0];
    a.active = inst_counter(ctx, 0, true);
    a.short_cells = inst_short(ctx, 0, true);
    inst_mark(ctx, 0, true);
    launch_coef(a, compute_dt, ctx->stream);
    inst_mark(ctx, 0, false);
    if (compute_dt) {
      a.dt_clears_count = 1;
      launch_dt(a, ctx->ninst, ctx->stream);
    }
    use_tiles(ctx, 1, a);
    a.zbytes = !ctx->bytes_off && SHAPE_TX[ctx->sel[1]] == ctx->mk_tx[0];
    a.active = inst_counter(ctx, 1, true);
    a.short_cells = inst_short(ctx, 1, true);
    inst_mark(ctx, 1, true);
    launch_update(a, ctx->stream);
    inst_mark(ctx, 1, false);
  } else if (ctx->es == 8) {
    enqueue_step_m<double>(ctx, in, out, s, compute_dt);
  } else {
    enqueue_step_m<float>(ctx, in, out, s, compute_dt);
  }
}

// Unext: U^{s-1} when the sweep's next reverse step reads it from a buffer
// that already holds it (the leaf path then makes its dual kact in this
// step's gather launch), else null.
void enqueue_adjoint(tg_ctx* ctx, void* const U[3], void* const lnext[3], void* const lout[3],
                     int s, bool dphi, void* const Unext[3] = nullptr) {
  const long long n = ctx->ncell;
  if (ctx->mode == TG_FP64) {
    AdjArgs a{};
    a.inst = ctx->inst_d;
    use_tiles(ctx, 2, a);
    a.U = Fields{(double*)U[0], (double*)U[1], (double*)U[2]};
    a.S = Statics{at<double>(ctx->statics, 0, n), at<double>(ctx->statics, 1, n),
                  at<double>(ctx->statics, 2, n), at<double>(ctx->statics, 3, n),
                  at<double>(ctx->statics, 4, n), at<double>(ctx->statics, 5, n),
                  at<double>(ctx->statics, 6, n)};
    a.lam_next = Fields{(double*)lnext[0], (double*)lnext[1], (double*)lnext[2]};
    a.dt_rec = ctx->dt_rec; a.max_steps = ctx->cfg.max_steps; a.step = s;
    a.ft = frict_tiny(ctx, s); a.tiny = ctx->cfg.tiny; a.gamma = ctx->cfg.gamma;
    a.lam_h = (double*)lout[0]; a.lam_hu = (double*)lout[1]; a.lam_hv = (double*)lout[2];
    a.dphi_part = dphi ? ctx->dphi_part : nullptr;
    a.tile_list = ctx->act_list;
    a.tile_count = ctx->act_count;
    a.mk_U = marks_of(ctx, U);
    a.mk_lnext = marks_of(ctx, lnext);
    a.mk_out = marks_of(ctx, lout);
    a.u_bytes = !ctx->bytes_off && SHAPE_TX[ctx->sel[2]] == ctx->mk_tx[0];
    a.l_bytes = !ctx->bytes_off && SHAPE_TX[ctx->sel[2]] == ctx->mk_tx[1];
    a.active = inst_counter(ctx, 2, true);
    a.short_cells = inst_short(ctx, 2, true);
    inst_mark(ctx, 2, true);
    if (ctx->leaves) {
      // the dual kact of U^s was made by the previous step's gather launch
      // when that launch was given this step's buffer
      const bool have_kd = ctx->kd_step == s && ctx->kd_U == U[0] && ctx->kd_dk == dphi;
      AdjArgs an = a;
      if (Unext) {
        an.U = Fields{(double*)Unext[0], (double*)Unext[1], (double*)Unext[2]};
        an.mk_U = marks_of(ctx, Unext);
        an.step = s - 1;
      }
      launch_leaf_adjoint(a, Unext ? &an : nullptr, !have_kd, ctx->leaf, ctx->lb, ctx->stream);
      ctx->kd_step = Unext ? s - 1 : -1;
      ctx->kd_U = Unext ? Unext[0] : nullptr;
      ctx->kd_dk = dphi;
    } else {
      launch_adjoint(a, ctx->stream);
    }
    inst_mark(ctx, 2, false);
  } else if (ctx->es == 8) {
    enqueue_adjoint_m<double>(ctx, U, lnext, lout, s, dphi);
  } else {
    enqueue_adjoint_m<float>(ctx, U, lnext, lout, s, dphi);
  }
}

int check_err(tg_ctx* ctx) {
  int e = 0;
  HIPCHK(ctx, hipMemcpyAsync(&e, ctx->err_d, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (e & 1) return set_err(ctx, TG_ERR_NONFINITE, "non-finite state (residual.C:137-147)");
  if (e & 2) return set_err(ctx, TG_ERR_EIGEN, "eigenvalue > 1e9 (get_coef_and_eigen.C:144)");
  return TG_OK;
}

// FP64 host array of one field of one member <-> device storage
int h2d_field(tg_ctx* ctx, void* dev_base, long long off, long long count, const double* src,
              int es = 0) {
  if ((es ? es : ctx->es) == 8)
    HIPCHK(ctx, hipMemcpyAsync((double*)dev_base + off, src, sizeof(double) * count,
                               hipMemcpyHostToDevice, ctx->stream));
  else {
    std::vector<float> tmp(count);
    for (long long i = 0; i < count; i++) tmp[i] = (float)src[i];
    HIPCHK(ctx, hipMemcpyAsync((float*)dev_base + off, tmp.data(), sizeof(float) * count,
                               hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));  // tmp goes out of scope
  }
  return TG_OK;
}

int d2h_field(tg_ctx* ctx, const void* dev_base, long long off, long long count, double* dst,
              int es = 0) {
  if ((es ? es : ctx->es) == 8) {
    HIPCHK(ctx, hipMemcpyAsync(dst, (const double*)dev_base + off, sizeof(double) * count,
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  } else {
    std::vector<float> tmp(count);
    HIPCHK(ctx, hipMemcpyAsync(tmp.data(), (const float*)dev_base + off, sizeof(float) * count,
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (long long i = 0; i < count; i++) dst[i] = (double)tmp[i];
  }
  return TG_OK;
}

int fill_stats(tg_ctx* ctx, tg_stats* st, int nst) {
  if (!st) return TG_OK;
  if (ctx->leaves)
    launch_leaf_stats(ctx->leaf,
                      Fields{(double*)ctx->U[0], (double*)ctx->U[1], (double*)ctx->U[2]},
                      ctx->part_mass, ctx->part_hmax, ctx->cfg.tiny, ctx->stream);
  else if (ctx->es == 8)
    launch_stats(ctx->inst_d, ctx->tiles_d[0], ctx->nt[0],
                 Fields{(double*)ctx->U[0], (double*)ctx->U[1], (double*)ctx->U[2]},
                 marks_of(ctx, ctx->U), ctx->part_mass, ctx->part_hmax, ctx->cfg.tiny,
                 ctx->stream);
  else
    launch_stats32(ctx->inst_d, ctx->tiles_d[0], ctx->nt[0],
                   FieldsT<float>{(float*)ctx->U[0], (float*)ctx->U[1], (float*)ctx->U[2]},
                   ctx->part_mass, ctx->part_hmax, ctx->cfg.tiny, ctx->stream);
  // mass, hmax, time and last dt of every member in one launch and one copy
  const int nm = ctx->ninst;
  const bool dt = ctx->rec_steps > 0;
  launch_member_reduce(ctx->part_mass, ctx->part_hmax, ctx->red_off[0], ctx->red_idx[0], nm,
                       ctx->time_d, dt ? ctx->dt_rec : nullptr, ctx->cfg.max_steps,
                       ctx->rec_steps - 1, ctx->red_out, ctx->stream);
  HIPCHK(ctx, hipMemcpyAsync(ctx->red_host, ctx->red_out, sizeof(double) * 4 * nm,
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  const double* r = ctx->red_host;
  for (int b = 0; b < nst && b < nm; b++) {
    st[b].dt = dt ? r[3 * nm + b] : 0.;
    st[b].time = r[2 * nm + b];
    st[b].mass = r[b];
    st[b].hmax = r[nm + b];
    st[b].step = ctx->rec_steps;
  }
  return TG_OK;
}

// Each member's dJ/dphi_bed from the adjoint's per-tile partials (tiles of
// the adjoint's variant), reduced on the device.
int member_dphi(tg_ctx* ctx, double* out) {
  const int nm = ctx->ninst, v = ctx->sel[2];
  launch_member_reduce(ctx->dphi_part, nullptr, ctx->red_off[v], ctx->red_idx[v], nm, nullptr,
                       nullptr, 0, 0, ctx->red_out, ctx->stream);
  HIPCHK(ctx, hipMemcpyAsync(ctx->red_host, ctx->red_out, sizeof(double) * nm,
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int b = 0; b < nm; b++) out[b] = ctx->red_host[b];
  return TG_OK;
}

int ensure_store(tg_ctx* ctx, std::vector<void*>& v, size_t count) {
  while (v.size() < count) {
    void* p = nullptr;
    int rc = dalloc(ctx, &p, ctx->state_bytes);
    if (rc) return rc;
    clear_marks(ctx, p);
    v.push_back(p);
  }
  return TG_OK;
}

// The stored trajectory lives in one allocation (fast to size per run);
// growing it frees the old block, so captured sweeps that point into it are
// dropped.
int ensure_traj(tg_ctx* ctx, size_t count) {
  if (count <= ctx->traj_cap) return TG_OK;
  drop_graph(ctx->gf);
  drop_graph(ctx->ga);
  drop_segment_graphs(ctx);
  dfree(ctx->traj_block);
  ctx->traj_cap = 0;
  int rc = dalloc(ctx, &ctx->traj_block, ctx->state_bytes * count);
  if (rc) return rc;
  ctx->traj_cap = count;
  for (size_t i = 0; i < count; i++)
    clear_marks(ctx, (char*)ctx->traj_block + i * ctx->state_bytes);
  return TG_OK;
}

void* traj_state(tg_ctx* ctx, int i) {
  return (char*)ctx->traj_block + (size_t)i * ctx->state_bytes;
}

void split(tg_ctx* ctx, void* base, void* out[3]) {
  for (int k = 0; k < 3; k++) out[k] = (char*)base + (size_t)k * ctx->es * ctx->ncell;
}

// copies the three fields and the zero marks
void copy_state(tg_ctx* ctx, void* dst, void* const src[3]) {
  const size_t b = (size_t)ctx->es * ctx->ncell;
  for (int k = 0; k < 3; k++) launch_copy(src[k], (char*)dst + k * b, b, ctx->stream);
  if (ctx->nsuper > 0)
    launch_copy((char*)src[0] + ctx->mk_off, (char*)dst + ctx->mk_off,
                sizeof(uint32_t) * (size_t)ctx->nsuper, ctx->stream);
}

// copy_state for a forward state that also gives the source and the copy
// the zero marks of their values (the update's tile shape), so that a state
// set from the host, whose marks say nothing, and its checkpoint are
// skipped where empty from the first step on; a plain copy where the
// uniform blocks are not there (leaf meshes, no marks)
void copy_state_marking(tg_ctx* ctx, void* dst, void* const src[3]) {
  const int v = ctx->sel[1];
  if (ctx->leaves || !ctx->marks || ctx->nb[v] != ctx->nsuper) {
    copy_state(ctx, dst, src);
    return;
  }
  launch_copy_marking(ctx->blocks_d[v], ctx->nb[v], v, ctx->es, src, dst, ctx->ncell,
                      marks_of(ctx, src), (uint32_t*)((char*)dst + ctx->mk_off), ctx->stream);
}

// The MaxH record of the state U just produced by step s.
void enqueue_box_max(tg_ctx* ctx, void* const U[3], int s) {
  if (ctx->nbox_chunks <= 0) return;
  if (fwd64(ctx) || ctx->es == 8)
    launch_box_max<double>(ctx->box_chunks, ctx->nbox_chunks, ctx->box_cells,
                           (const double*)U[0], ctx->box_rec, ctx->cfg.max_steps, s, ctx->stream);
  else
    launch_box_max<float>(ctx->box_chunks, ctx->nbox_chunks, ctx->box_cells, (const float*)U[0],
                          ctx->box_rec, ctx->cfg.max_steps, s, ctx->stream);
}
void zero_box_record(tg_ctx* ctx) {
  if (ctx->nbox_chunks > 0)
    launch_zero(ctx->box_rec, sizeof(double) * ctx->ninst * (size_t)ctx->cfg.max_steps,
                ctx->stream);
}

// Copies (src) or zeroes (src null) member b's cells of a state / adjoint
// buffer; a copy clears the member's zero marks.
void member_fill(tg_ctx* ctx, const void* src, void* dst, int b) {
  const InstDev& I = ctx->inst_h[b];
  long long mw0 = 0, nmw = 0;
  if (ctx->marks) {
    mw0 = ctx->leaves ? 0 : I.soff;
    nmw = ctx->leaves ? ctx->nsuper : (long long)I.nsx * I.nsy;
  }
  launch_member_fill(src, dst, ctx->es, ctx->ncell, I.off, (long long)I.nx * I.ny,
                     ctx->marks ? (uint32_t*)((char*)dst + ctx->mk_off) : nullptr, mw0, nmw,
                     ctx->stream);
}

bool same_key(const SweepGraph& G, const long long key[8]) {
  if (!G.valid) return false;
  for (int i = 0; i < 8; i++)
    if (G.key[i] != key[i]) return false;
  return true;
}

// Captures what `body` enqueues on the context's stream into G (replacing
// what G held) and instantiates it.
template <class F>
int capture(tg_ctx* ctx, SweepGraph& G, const long long key[8], F&& body) {
  drop_graph(G);
  HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  body();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (e != hipSuccess || !g)
    return set_err(ctx, TG_ERR_HIP, std::string("stream capture failed: ") + hipGetErrorString(e));
  G.g = g;
  HIPCHK(ctx, hipGraphInstantiate(&G.x, G.g, nullptr, nullptr, 0));
  for (int i = 0; i < 8; i++) G.key[i] = key[i];
  G.valid = true;
  return TG_OK;
}

// Where each state of a forward sweep lives: U^s for s in [s0, N) is kept
// in traj[s - s0] (written there directly by the step that produces it);
// below s0 the states alternate between the two working buffers and U^s is
// checkpointed every k steps.  U^N ends in a working buffer.
int stored_steps(const tg_ctx* ctx, int nsteps) {
  const double state = (double)ctx->state_bytes;
  const long long m = (long long)((double)ctx->cfg.store_bytes / state);
  return (int)(m < nsteps ? (m < 0 ? 0 : m) : nsteps);
}

int forward_loop(tg_ctx* ctx, const tg_params* prm, int nsteps, const double* dt_seq,
                 tg_stats* stats, int nstats) {
  if (!ctx->have_mesh || !ctx->have_state) return set_err(ctx, TG_ERR_STATE, "mesh/state not set");
  if (nsteps < 0 || nsteps > ctx->cfg.max_steps)
    return set_err(ctx, TG_ERR_ARG, "nsteps outside [0, max_steps]");
  int rc = upload_params(ctx, prm);
  if (rc) return rc;
  const int k = ctx->cfg.checkpoint_every;
  int M = stored_steps(ctx, nsteps);
  if (ensure_traj(ctx, (size_t)M) != TG_OK) {  // no room: checkpoints only
    M = 0;
    (void)hipGetLastError();  // the handled allocation failure is not the run's error
    ctx->err.clear();
  }
  const int s0 = nsteps - M;
  rc = ensure_store(ctx, ctx->ckpt, (size_t)(s0 + k - 1) / k + 1);
  if (rc) return rc;
  ctx->rec_steps = ctx->run_steps = 0;
  if (dt_seq)  // the same prescribed sequence for every member
    for (int b = 0; b < ctx->ninst; b++)
      HIPCHK(ctx, hipMemcpyAsync(ctx->dt_rec + (long long)b * ctx->cfg.max_steps, dt_seq,
                                 sizeof(double) * nsteps, hipMemcpyHostToDevice, ctx->stream));
  // the sweep: U^s -> U^{s+1} for s = 0..nsteps-1 (pointer roles planned up front)
  void* A[3] = {ctx->U[0], ctx->U[1], ctx->U[2]};
  void* B[3] = {ctx->Ut[0], ctx->Ut[1], ctx->Ut[2]};
  auto body = [&]() -> void** {
    launch_reset_red(ctx->red, ctx->ninst, ctx->stream);
    zero_box_record(ctx);
    void** cur = A;
    void* slot[3];
    for (int s = 0; s < nsteps; s++) {
      if (s < s0 && s % k == 0) copy_state_marking(ctx, ctx->ckpt[s / k], cur);
      if (s == s0 && s0 == 0 && M > 0) copy_state_marking(ctx, traj_state(ctx, 0), cur);
      void* dst[3];
      void** dstp;
      if (s + 1 >= s0 && s + 1 < nsteps) {  // U^{s+1} is kept: write it in place
        split(ctx, traj_state(ctx, s + 1 - s0), dst);
        dstp = dst;
      } else {
        dstp = (cur == A) ? B : A;
      }
      enqueue_step(ctx, cur, dstp, s, dt_seq == nullptr);
      enqueue_box_max(ctx, dstp, s);
      if (dstp == dst) {
        for (int c = 0; c < 3; c++) slot[c] = dst[c];
        cur = slot;
      } else {
        cur = dstp;
      }
    }
    return cur;  // A or B (U^nsteps never goes to the stored trajectory)
  };
  void** endU = nullptr;
  if (ctx->cfg.use_graph && !ctx->instrument) {
    const long long key[8] = {nsteps, (dt_seq != nullptr) + 2 * frict_switch_local(ctx),
                              (long long)(intptr_t)ctx->U[0], ctx->sel[0], ctx->sel[1], k, M,
                              ctx->nbox_chunks > 0 ? ctx->box_version : 0};
    if (!same_key(ctx->gf, key)) {
      rc = capture(ctx, ctx->gf, key, [&]() { endU = body(); });
      if (rc) return rc;
      ctx->gf_end_is_b = (endU == B);
    }
    endU = ctx->gf_end_is_b ? B : A;
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    HIPCHK(ctx, hipGraphLaunch(ctx->gf.x, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  } else {
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    endU = body();
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  }
  if (endU == B)
    for (int c = 0; c < 3; c++) std::swap(ctx->U[c], ctx->Ut[c]);
  HIPCHK(ctx, hipGetLastError());
  rc = check_err(ctx);
  if (rc) return rc;
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  ctx->fwd_ms = ms;
  ctx->rec_steps = ctx->run_steps = nsteps;
  ctx->traj_s0 = s0;
  if (dt_seq) {  // time advances by the prescribed sequence
    std::vector<double> t(ctx->ninst);
    HIPCHK(ctx, hipMemcpy(t.data(), ctx->time_d, sizeof(double) * ctx->ninst,
                          hipMemcpyDeviceToHost));
    for (auto& v : t)
      for (int s = 0; s < nsteps; s++) v += dt_seq[s];
    HIPCHK(ctx, hipMemcpy(ctx->time_d, t.data(), sizeof(double) * ctx->ninst,
                          hipMemcpyHostToDevice));
  }
  return fill_stats(ctx, stats, nstats);
}

// FNV-1a over 64-bit words (graph keys of staged functionals)
long long fnv(const std::vector<long long>& v) {
  unsigned long long h = 1469598103934665603ull;
  for (long long x : v)
    for (int i = 0; i < 8; i++) {
      h ^= (unsigned long long)((x >> (8 * i)) & 0xff);
      h *= 1099511628211ull;
    }
  return (long long)(h >> 1);
}

// Reverse sweep.  steps[b] (null: every member at N) is the state U^{steps[b]}
// that member b's weights apply to: lambda^{steps[b]} = w_b, lambda^s = 0
// above it (calc_adjoint.C:49 sets func_sens at the MaxH step only, zero at
// the others).  The sweep starts at S = max_b steps[b].
int adjoint_loop(tg_ctx* ctx, const tg_params* prm, const double* w, const int* steps,
                 double* lambda0, double* dJ) {
  if (ctx->run_steps <= 0) return set_err(ctx, TG_ERR_STATE, "no forward run recorded");
  ctx->kd_step = -1;  // nothing made ahead for this sweep
  const int N = ctx->run_steps, k = ctx->cfg.checkpoint_every, s0 = ctx->traj_s0;
  std::vector<int> st(ctx->ninst, N);
  int S = 0;
  for (int b = 0; b < ctx->ninst; b++) {
    if (steps) {
      if (steps[b] < 0 || steps[b] > N)
        return set_err(ctx, TG_ERR_ARG,
                       "functional step of member " + std::to_string(b) +
                           " outside [0, steps of the last forward run]");
      st[b] = steps[b];
    }
    S = st[b] > S ? st[b] : S;
  }
  bool staged = false;  // some member's weights enter below S
  for (int b = 0; b < ctx->ninst; b++) staged |= st[b] < S;
  int rc = upload_params(ctx, prm);
  if (rc) return rc;
  const long long n = ctx->ncell;
  rc = ensure_store(ctx, ctx->seg, (size_t)(s0 > 0 ? k : 1));
  if (rc) return rc;
  const bool dphi = dJ != nullptr;
  // lambda^S = func_sens (calc_adjoint.C:93-97): from the host, or (w null)
  // the weights kept on the device; members whose step lies below S start
  // from zero and take their weights when the sweep reaches their step
  const void* src = nullptr;
  if (w && staged) {
    if (!ctx->wsrc) {
      rc = dalloc(ctx, &ctx->wsrc, ctx->state_bytes);
      if (rc) return rc;
    }
    void* f[3];
    split(ctx, ctx->wsrc, f);
    for (int c = 0; c < 3; c++) {
      rc = h2d_field(ctx, f[c], 0, n, w + c * n);
      if (rc) return rc;
    }
    clear_marks(ctx, ctx->wsrc);
    copy_state(ctx, ctx->lam[0][0], f);
    src = ctx->wsrc;
  } else if (w) {
    for (int c = 0; c < 3; c++) {
      rc = h2d_field(ctx, ctx->lam[0][c], 0, n, w + c * n);
      if (rc) return rc;
    }
    clear_marks(ctx, ctx->lam[0][0]);
  } else {
    src = ctx->fw;
  }
  // the weights kept on the device are read in place by the first reverse
  // step (lambda^S) unless members enter later (lambda^S then differs from
  // them) or there is no step (lambda^0 = lambda^S is returned)
  void* top[3] = {nullptr, nullptr, nullptr};
  if (!w && !staged && S > 0) split(ctx, ctx->fw, top);
  else if (!w) {
    void* f[3];
    split(ctx, ctx->fw, f);
    copy_state(ctx, ctx->lam[0][0], f);
  }
  auto lam_next = [&](int s, int cur) -> void* const* {
    return (top[0] && s == S - 1) ? top : ctx->lam[cur];
  };
  std::vector<std::vector<int>> enter(S);  // members whose weights enter at lambda^s
  for (int b = 0; b < ctx->ninst; b++) {
    if (st[b] < S) {
      member_fill(ctx, nullptr, ctx->lam[0][0], b);
      enter[st[b]].push_back(b);
    }
  }
  auto inject = [&](int s, void* lout) {
    for (int b : enter[s]) member_fill(ctx, src, lout, b);
  };
  // steps S-1 .. s0 read the stored trajectory; below s0 the sweep works in
  // segments of k: restore U^{a} from its checkpoint, recompute U^{a+1..b-1}
  // with the recorded dt, then the adjoint steps b-1 .. a (segments at or
  // above S are not needed, the one holding S stops below it)
  const int nseg = (s0 + k - 1) / k;
  const bool need_gb = nseg > 0 && S > 0;
  const int nstored = S > s0 ? S - s0 : 0;
  auto seg_end = [&](int a) {
    const int b = (a + k < s0) ? a + k : s0;
    return b < S ? b : S;
  };
  auto stored = [&](int& cur) {
    for (int s = S - 1; s >= s0; s--) {
      void* Us[3];
      void* Un[3];
      split(ctx, traj_state(ctx, s - s0), Us);
      const bool next = s - 1 >= s0;  // U^{s-1} stored too
      if (next) split(ctx, traj_state(ctx, s - 1 - s0), Un);
      enqueue_adjoint(ctx, Us, lam_next(s, cur), ctx->lam[1 - cur], s, dphi,
                      next ? Un : nullptr);
      inject(s, ctx->lam[1 - cur][0]);
      cur = 1 - cur;
    }
  };
  // segment sg's states: U^a read in place from its checkpoint (no kernel
  // writes a step's input), U^{a+1} .. U^{b-1} recomputed into seg[1 ..]
  auto seg_state = [&](int sg, int s, void* out[3]) {
    const int a = sg * k;
    split(ctx, s == a ? ctx->ckpt[sg] : ctx->seg[s - a], out);
  };
  auto recompute = [&](int sg) {
    const int a = sg * k, b = seg_end(a);
    for (int s = a; s + 1 < b; s++) {
      void* in[3];
      void* out[3];
      seg_state(sg, s, in);
      seg_state(sg, s + 1, out);
      enqueue_step(ctx, in, out, s, false);
    }
  };
  auto reverse = [&](int sg, int& cur) {
    const int a = sg * k, b = seg_end(a);
    for (int s = b - 1; s >= a; s--) {
      void* Us[3];
      void* Un[3];
      seg_state(sg, s, Us);
      const bool next = s - 1 >= a;  // U^{s-1} recomputed in this segment
      if (next) seg_state(sg, s - 1, Un);
      enqueue_adjoint(ctx, Us, lam_next(s, cur), ctx->lam[1 - cur], s, dphi,
                      next ? Un : nullptr);
      inject(s, ctx->lam[1 - cur][0]);
      cur = 1 - cur;
    }
  };
  int cur = 0;
  float rec_ms = 0, ms = 0;
  // timing: ev[2] .. stored part .. ev[6] .. segments .. ev[3]; the i-th
  // visited segment's recompute lies between seg_ev[2i] and seg_ev[2i + 1]
  int nvisit = 0;
  for (int sg = nseg - 1; sg >= 0; sg--)
    if (sg * k < S) nvisit++;
  while ((int)ctx->seg_ev.size() < 2 * nvisit) {
    hipEvent_t e = nullptr;
    HIPCHK(ctx, hipEventCreate(&e));
    ctx->seg_ev.push_back(e);
  }
  const bool graph = ctx->cfg.use_graph && !ctx->instrument;
  if (graph) {
    long long stage = 0;  // what the captured launches depend on beyond N
    if (S < N || staged) {
      std::vector<long long> v(st.begin(), st.end());
      v.push_back(S);
      v.push_back((long long)(intptr_t)src);
      stage = fnv(v);
    }
    const long long key[8] = {N, k + (top[0] ? (1LL << 32) : 0),
                              dphi + 2 * frict_switch_local(ctx), ctx->sel[0], ctx->sel[1],
                              ctx->sel[2], s0, stage};
    if (!same_key(ctx->ga, key)) {
      drop_segment_graphs(ctx);
      rc = capture(ctx, ctx->ga, key, [&]() {
        int c = 0;
        if (dphi) launch_zero(ctx->dphi_part, sizeof(double) * ctx->ntiles, ctx->stream);
        else launch_zero(ctx->act_count, sizeof(int), ctx->stream);  // the graph is never empty
        stored(c);
      });
      if (rc) return rc;
      ctx->gr.assign(nseg, SweepGraph{});
      ctx->gs.assign(nseg, SweepGraph{});
      int c = nstored % 2;
      for (int sg = nseg - 1; sg >= 0 && need_gb; sg--) {
        if (sg * k >= S) continue;
        rc = capture(ctx, ctx->gr[sg], key, [&]() { recompute(sg); });
        if (rc) return rc;
        rc = capture(ctx, ctx->gs[sg], key, [&]() { reverse(sg, c); });
        if (rc) return rc;
      }
    }
  } else if (dphi) {
    launch_zero(ctx->dphi_part, sizeof(double) * ctx->ntiles, ctx->stream);
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
  if (graph) HIPCHK(ctx, hipGraphLaunch(ctx->ga.x, ctx->stream));
  else stored(cur);
  HIPCHK(ctx, hipEventRecord(ctx->ev[6], ctx->stream));
  int iv = 0;
  for (int sg = nseg - 1; sg >= 0 && need_gb; sg--) {
    if (sg * k >= S) continue;
    HIPCHK(ctx, hipEventRecord(ctx->seg_ev[2 * iv], ctx->stream));
    if (graph) HIPCHK(ctx, hipGraphLaunch(ctx->gr[sg].x, ctx->stream));
    else recompute(sg);
    HIPCHK(ctx, hipEventRecord(ctx->seg_ev[2 * iv + 1], ctx->stream));
    if (graph) HIPCHK(ctx, hipGraphLaunch(ctx->gs[sg].x, ctx->stream));
    else reverse(sg, cur);
    iv++;
  }
  if (graph) cur = S % 2;
  HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
  HIPCHK(ctx, hipGetLastError());
  rc = check_err(ctx);
  if (rc) return rc;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
  ctx->adj_ms = ms;
  for (int i = 0; i < iv; i++) {
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->seg_ev[2 * i], ctx->seg_ev[2 * i + 1]));
    rec_ms += ms;
  }
  ctx->rec_ms = rec_ms;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[6]));
  ctx->stored_ms = ms;
  ctx->stored_steps = nstored;
  if (lambda0)
    for (int c = 0; c < 3; c++) {
      rc = d2h_field(ctx, ctx->lam[cur][c], 0, n, lambda0 + c * n);
      if (rc) return rc;
    }
  ctx->member_dphi.assign(ctx->ninst, 0.);
  if (dphi) {
    rc = member_dphi(ctx, ctx->member_dphi.data());
    if (rc) return rc;
    double acc = 0;
    for (double v : ctx->member_dphi) acc += v;
    *dJ = acc;
  }
  return TG_OK;
}

// Builds members, tiles and device buffers.
// Reorders a row-major tile list so that workgroup b (which runs on XCD
// b % NXCD) takes tile start[x] + b / NXCD of band x = b % NXCD: band x is a
// contiguous run of the list, of q or q+1 tiles (q = n / NXCD, the first
// n % NXCD bands one longer).  A bijection on [0, n).
void band_by_xcd(std::vector<TileDev>& t) {
  const int n = (int)t.size(), q = n / NXCD, r = n % NXCD;
  std::vector<TileDev> out(n);
  for (int b = 0; b < n; b++) {
    const int x = b % NXCD, slot = b / NXCD;
    const int start = x * q + (x < r ? x : r);
    out[b] = t[start + slot];
  }
  t.swap(out);
}

// Checks the leaf tables of a TG_MESH_LEAVES mesh and builds the device
// tables: slot-major neighbours and the per-leaf meta word (side kinds,
// which_neighbor back slots).
int leaf_tables(tg_ctx* ctx, const tg_mesh* m, std::vector<int>& nbrT,
                std::vector<uint32_t>& meta) {
  const int n = m->n;
  if (n < 1 || !m->nbr || !m->nproc || !m->gen || !m->cdx || !m->xc || !m->grav || !m->dgrav ||
      !m->curv)
    return set_err(ctx, TG_ERR_ARG, "bad leaf mesh");
  nbrT.assign((size_t)8 * n, -1);
  meta.assign(n, 0u);
  const int32_t* nb = m->nbr;
  const int32_t* np = m->nproc;
  auto bad = [&](int e, const char* why) {
    return set_err(ctx, TG_ERR_ARG, std::string("leaf ") + std::to_string(e) + ": " + why);
  };
  for (int e = 0; e < n; e++) {
    if (!(m->cdx[2 * e] > 0) || !(m->cdx[2 * e + 1] > 0)) return bad(e, "cell size <= 0");
    uint32_t kinds = 0;
    for (int s = 0; s < 4; s++) {
      const int p1 = np[8 * e + s], p2 = np[8 * e + s + 4];
      const int f1 = nb[8 * e + s], f2 = nb[8 * e + s + 4];
      uint32_t k;
      if (p1 == -1) {
        if (p2 >= 0) return bad(e, "second neighbour on a boundary side");
        k = LEAF_BOUNDARY;
      } else if (p1 >= 0) {
        if (f1 < 0 || f1 >= n || f1 == e) return bad(e, "neighbour index out of range");
        if (p2 >= 0) {
          if (f2 < 0 || f2 >= n || f2 == e || f2 == f1)
            return bad(e, "second neighbour index out of range");
          if (m->gen[f1] != m->gen[e] + 1 || m->gen[f2] != m->gen[e] + 1)
            return bad(e, "two neighbours on a side must be one generation finer");
          k = LEAF_FINER;
          nbrT[(size_t)(s + 4) * n + e] = f2;
        } else if (p2 == -2) {
          if (m->gen[f1] == m->gen[e]) k = LEAF_SAME;
          else if (m->gen[f1] == m->gen[e] - 1) k = LEAF_COARSER;
          else return bad(e, "single neighbour more than one generation apart");
        } else {
          return bad(e, "bad neigh_proc of a second slot");
        }
        nbrT[(size_t)s * n + e] = f1;
      } else {
        return bad(e, "bad neigh_proc");
      }
      kinds |= k << (2 * s);
    }
    if ((kinds & 3u) == LEAF_BOUNDARY && ((kinds >> 4) & 3u) == LEAF_BOUNDARY)
      return bad(e, "domain boundary on both y sides");
    if (((kinds >> 2) & 3u) == LEAF_BOUNDARY && ((kinds >> 6) & 3u) == LEAF_BOUNDARY)
      return bad(e, "domain boundary on both x sides");
    meta[e] = kinds;
  }
  // back slots: where each present neighbour lists this leaf
  for (int e = 0; e < n; e++)
    for (int k = 0; k < 8; k++) {
      const uint32_t ks = (meta[e] >> (2 * (k & 3))) & 3u;
      const bool present = (k < 4) ? ks != LEAF_BOUNDARY : ks == LEAF_FINER;
      if (!present) continue;
      const int f = nbrT[(size_t)k * n + e];
      int w = -1;
      for (int i = 0; i < 8 && w < 0; i++)
        if (nb[8 * f + i] == e && np[8 * f + i] >= 0) w = i;
      if (w < 0) return bad(e, "a neighbour does not list the leaf back");
      meta[e] |= (uint32_t)w << (8 + 3 * k);
    }
  return TG_OK;
}

long long cells_of(const tg_mesh* m) {
  return m->kind == TG_MESH_LEAVES ? (long long)m->n : (long long)m->nx * m->ny;
}

int set_members(tg_ctx* ctx, int nm, const tg_mesh* meshes, const tg_params* params) {
  if (nm < 1 || !meshes) return set_err(ctx, TG_ERR_ARG, "no members");
  const bool leaves = meshes[0].kind == TG_MESH_LEAVES;
  std::vector<int> nbrT;
  std::vector<uint32_t> meta;
  if (leaves) {
    if (nm != 1) return set_err(ctx, TG_ERR_UNSUPPORTED, "TG_MESH_LEAVES takes one member");
    if (ctx->mode != TG_FP64)
      return set_err(ctx, TG_ERR_UNSUPPORTED, "TG_MESH_LEAVES is built for TG_FP64");
    int rc = leaf_tables(ctx, &meshes[0], nbrT, meta);
    if (rc) return rc;
  }
  for (int b = 0; b < nm && !leaves; b++) {
    const tg_mesh* m = &meshes[b];
    if (m->kind != TG_MESH_UNIFORM)
      return set_err(ctx, TG_ERR_UNSUPPORTED, "members must all be TG_MESH_UNIFORM");
    if (m->nx < 1 || m->ny < 1 || !(m->dx > 0) || !(m->dy > 0) || !m->grav || !m->dgrav ||
        !m->curv)
      return set_err(ctx, TG_ERR_ARG, "bad mesh");
    // The kernels use dx, dy as the centre spacing; the reference takes
    // differences of centre coordinates (element2.C:894-895).  Both agree
    // only if those differences are exact (dyadic spacings and origins).
    for (int i = 0; i + 1 < m->nx; i++)
      if ((m->x0 + (i + 1) * m->dx) - (m->x0 + i * m->dx) != m->dx)
        return set_err(ctx, TG_ERR_ARG, "x centre spacing is not exactly dx");
    for (int j = 0; j + 1 < m->ny; j++)
      if ((m->y0 + (j + 1) * m->dy) - (m->y0 + j * m->dy) != m->dy)
        return set_err(ctx, TG_ERR_ARG, "y centre spacing is not exactly dy");
  }
  for (int b = 0; b < nm; b++)
    if (meshes[b].mat)
      for (long long c = 0; c < cells_of(&meshes[b]); c++)
        if (meshes[b].mat[c] >= MAX_MAT)
          return set_err(ctx, TG_ERR_ARG, "material index of member " + std::to_string(b) +
                                              " cell " + std::to_string(c) + " >= " +
                                              std::to_string(MAX_MAT));
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  free_fields(ctx);
  ctx->have_mesh = ctx->have_state = false;
  ctx->rec_steps = ctx->run_steps = 0;
  ctx->ninst = nm;
  ctx->inst_h.assign(nm, InstDev{});
  for (auto& t : ctx->tiles_h) t.clear();
  ctx->nsuper = 0;
  long long off = 0;
  for (int b = 0; b < nm && leaves; b++) {  // one member, one thread per leaf
    InstDev& I = ctx->inst_h[0];
    I.nx = meshes[0].n; I.ny = 1; I.off = 0;
    I.maxtime = I.nextoutput = I.nextsave = HUGE_VAL;
    I.mat_base = 0;
    if (params) params_into(I, &params[0]);
    for (int v = 0; v < NSHAPE; v++)
      for (int x = 0; x < meshes[0].n; x += BLOCK)
        ctx->tiles_h[v].push_back(TileDev{0, x, 0, 0, -1, -1, -1, -1});
    off = meshes[0].n;
  }
  // leaf chunks (the BLOCK leaves of one workgroup) carry one zero-mark word
  // each; a chunk's neighbour chunks are those holding its leaves' neighbours
  std::vector<int> cnb_off(1, 0), cnb;
  if (leaves) {
    const int n = meshes[0].n, nq = (n + BLOCK - 1) / BLOCK;
    ctx->nsuper = nq;
    std::vector<int> seen(nq, -1);
    for (int q = 0; q < nq; q++) {
      for (int e = q * BLOCK; e < n && e < (q + 1) * BLOCK; e++)
        for (int k = 0; k < 8; k++) {
          const uint32_t ks = (meta[e] >> (2 * (k & 3))) & 3u;
          const bool present = (k < 4) ? ks != LEAF_BOUNDARY : ks == LEAF_FINER;
          if (!present) continue;
          const int r = nbrT[(size_t)k * n + e] / BLOCK;
          if (r != q && seen[r] != q) {
            seen[r] = q;
            cnb.push_back(r);
          }
        }
      cnb_off.push_back((int)cnb.size());
    }
  }
  for (int b = 0; b < nm && !leaves; b++) {
    const tg_mesh* m = &meshes[b];
    InstDev& I = ctx->inst_h[b];
    I.nx = m->nx; I.ny = m->ny; I.off = off;
    I.nsx = (m->nx + MARK_W - 1) / MARK_W;
    I.nsy = (m->ny + MARK_H - 1) / MARK_H;
    I.soff = ctx->nsuper;
    ctx->nsuper += (long long)I.nsx * I.nsy;
    I.dx = m->dx; I.dy = m->dy;
    I.inv_dx = 1 / m->dx; I.inv_sx = 1 / (m->dx + m->dx);
    I.inv_dy = 1 / m->dy; I.inv_sy = 1 / (m->dy + m->dy);
    I.maxtime = I.nextoutput = I.nextsave = HUGE_VAL;
    I.mat_base = b * MAX_MAT;
    if (params) params_into(I, &params[b]);
    for (int v = 0; v < NSHAPE; v++) {
      const int TX = SHAPE_TX[v], TY = BLOCK / SHAPE_TX[v];
      for (int ty = 0; ty < m->ny; ty += TY)
        for (int tx = 0; tx < m->nx; tx += TX) {
          const int bx = tx / MARK_W, by = ty / MARK_H;
          const int mw = (int)I.soff + by * I.nsx + bx;
          ctx->tiles_h[v].push_back(TileDev{b, tx, ty, mw, bx > 0 ? mw - 1 : -1,
                                            bx + 1 < I.nsx ? mw + 1 : -1,
                                            by > 0 ? mw - I.nsx : -1,
                                            by + 1 < I.nsy ? mw + I.nsx : -1});
        }
    }
    off += (long long)m->nx * m->ny;
  }
  ctx->ncell = off;
  ctx->ntiles = 0;
  for (int v = 0; v < NSHAPE; v++) {
    if (SHAPE_XCD[v] && !leaves) band_by_xcd(ctx->tiles_h[v]);
    ctx->nt[v] = (int)ctx->tiles_h[v].size();
    if (ctx->nt[v] > ctx->ntiles) ctx->ntiles = ctx->nt[v];
    // mark blocks in the order their first tile appears in the table (a
    // banded table stays roughly banded); tile byte as in the kernels
    ctx->blocks_h[v].clear();
    if (leaves) continue;
    const int TX = SHAPE_TX[v], TY = BLOCK / SHAPE_TX[v];
    std::unordered_map<int, int> at;
    for (int i = 0; i < ctx->nt[v]; i++) {
      const TileDev& t = ctx->tiles_h[v][i];
      auto it = at.find(t.mw);
      if (it == at.end()) {
        it = at.emplace(t.mw, (int)ctx->blocks_h[v].size()).first;
        const InstDev& I = ctx->inst_h[t.inst];
        ctx->blocks_h[v].push_back(BlockDev{t.inst, t.mw, t.mwW, t.mwE, t.mwS, t.mwN,
                                            {-1, -1, -1, -1}, t.x0 - t.x0 % MARK_W,
                                            t.y0 - t.y0 % MARK_H, I.nx, I.ny, I.off});
      }
      const int byte = ((t.y0 % MARK_H) / TY) * (MARK_W / TX) + (t.x0 % MARK_W) / TX;
      ctx->blocks_h[v][it->second].tile[byte] = i;
    }
    // every mark word has one block: index the table by it, so a workgroup
    // reads its block's marks without waiting for the table entry
    std::vector<BlockDev> by_word(ctx->blocks_h[v].size());
    for (const BlockDev& B : ctx->blocks_h[v]) by_word[B.mw] = B;
    ctx->blocks_h[v].swap(by_word);
  }
  for (int v = 0; v < NSHAPE; v++) {
    ctx->nb[v] = (int)ctx->blocks_h[v].size();
  }
  const long long n = ctx->ncell;
  const size_t fb = (size_t)ctx->es * n;
  auto align256 = [](size_t b) { return (b + 255) / 256 * 256; };
  ctx->mk_off = align256(3 * fb);
  ctx->state_bytes = ctx->mk_off + align256(sizeof(uint32_t) * (size_t)ctx->nsuper);
  // every kernel of the uniform path keeps zero marks (FP64 and the FP32
  // kernels of the other precisions), and so do the leaf kernels (FP64
  // only), whose marks are per chunk of leaves
  ctx->marks = ctx->nsuper > 0;
  int rc = 0;
  void** bufs[4] = {ctx->U, ctx->Ut, ctx->lam[0], ctx->lam[1]};
  for (int i = 0; i < 4 && !rc; i++) {
    rc = dalloc(ctx, &bufs[i][0], ctx->state_bytes);
    if (!rc) {
      bufs[i][1] = (char*)bufs[i][0] + fb;
      bufs[i][2] = (char*)bufs[i][0] + 2 * fb;
      clear_marks(ctx, bufs[i][0]);
    }
  }
  const size_t sb = (size_t)ctx->ses * n;
  ctx->U_home = ctx->U[0];
  if (!rc) rc = dalloc(ctx, &ctx->statics, 7 * sb);
  if (!rc && ctx->mode == TG_FWD64_ADJ32)
    rc = dalloc(ctx, &ctx->statics_f, 7 * sizeof(float) * n);
  if (!rc) rc = dalloc(ctx, &ctx->kact, sb);
  if (!rc) rc = dalloc(ctx, (void**)&ctx->part_mass, sizeof(double) * ctx->ntiles);
  if (!rc) rc = dalloc(ctx, (void**)&ctx->part_hmax, sizeof(double) * ctx->ntiles);
  if (!rc) rc = dalloc(ctx, (void**)&ctx->red_out, sizeof(double) * 4 * nm);
  for (int v = 0; v < NSHAPE && !rc; v++) {
    rc = dalloc(ctx, (void**)&ctx->red_off[v], sizeof(int) * (nm + 1));
    if (!rc) rc = dalloc(ctx, (void**)&ctx->red_idx[v], sizeof(int) * (ctx->nt[v] + 1));
  }
  if (!rc && hipHostMalloc((void**)&ctx->red_host, sizeof(double) * 4 * nm,
                           hipHostMallocDefault) != hipSuccess)
    rc = set_err(ctx, TG_ERR_NOMEM, "pinned host buffer of the per-member results");
  if (!rc) rc = dalloc(ctx, (void**)&ctx->dphi_part, sizeof(double) * ctx->ntiles);
  if (!rc && ctx->mode != TG_FP64)
    rc = dalloc(ctx, (void**)&ctx->thin_list, sizeof(int2) * (size_t)n);
  if (!rc && ctx->mode != TG_FP64) rc = dalloc(ctx, (void**)&ctx->thin_count, sizeof(int));
  if (!rc) rc = dalloc(ctx, (void**)&ctx->act_list, sizeof(int) * (size_t)ctx->ntiles);
  if (!rc) rc = dalloc(ctx, (void**)&ctx->act_count, sizeof(int));
  if (!rc) rc = dalloc(ctx, (void**)&ctx->inst_d, sizeof(InstDev) * nm);
  for (int v = 0; v < NSHAPE && !rc; v++)
    rc = dalloc(ctx, (void**)&ctx->tiles_d[v], sizeof(TileDev) * ctx->nt[v]);
  for (int v = 0; v < NSHAPE && !rc; v++)
    if (ctx->nb[v] > 0)
      rc = dalloc(ctx, (void**)&ctx->blocks_d[v], sizeof(BlockDev) * ctx->nb[v]);
  bool any_mat = false;
  for (int b = 0; b < nm; b++) any_mat |= meshes[b].mat != nullptr;
  if (!rc && any_mat) rc = dalloc(ctx, (void**)&ctx->mat_d, (size_t)n);
  if (!rc && any_mat) rc = dalloc(ctx, (void**)&ctx->mtab_d, sizeof(MatDev) * MAX_MAT * nm);
  if (!rc) rc = dalloc(ctx, (void**)&ctx->red, sizeof(RedDev) * nm);
  if (!rc) rc = dalloc(ctx, (void**)&ctx->time_d, sizeof(double) * nm);
  if (!rc) rc = dalloc(ctx, (void**)&ctx->dt_rec, sizeof(double) * nm * ctx->cfg.max_steps);
  if (leaves) {
    if (!rc) rc = dalloc(ctx, (void**)&ctx->leaf_nbr, sizeof(int) * 8 * (size_t)n);
    if (!rc) rc = dalloc(ctx, (void**)&ctx->leaf_meta, sizeof(uint32_t) * (size_t)n);
    if (!rc) rc = dalloc(ctx, (void**)&ctx->leaf_cdx, sizeof(double) * 2 * (size_t)n);
    if (!rc) rc = dalloc(ctx, (void**)&ctx->leaf_xc, sizeof(double) * 2 * (size_t)n);
    if (!rc) rc = dalloc(ctx, (void**)&ctx->lb.kd, sizeof(double) * (size_t)n);
    if (!rc) rc = dalloc(ctx, (void**)&ctx->lb.dkd, sizeof(double) * (size_t)n);
    if (!rc) rc = dalloc(ctx, (void**)&ctx->lb.jac, sizeof(double) * 81 * (size_t)n);
    if (!rc) rc = dalloc(ctx, (void**)&ctx->lb.kdmk, sizeof(uint32_t) * (size_t)ctx->nsuper);
    if (!rc) rc = dalloc(ctx, (void**)&ctx->leaf_cnb_off, sizeof(int) * cnb_off.size());
    if (!rc) rc = dalloc(ctx, (void**)&ctx->leaf_cnb, sizeof(int) * (cnb.empty() ? 1 : cnb.size()));
  }
  if (rc) {
    free_fields(ctx);
    return rc;
  }
  // static coefficients: field k of member b at statics[k*n + off_b]
  for (int b = 0; b < nm && !rc; b++) {
    const tg_mesh* m = &meshes[b];
    const long long nb = cells_of(m), o = ctx->inst_h[b].off;
    const double* src[7] = {m->grav, m->grav + nb, m->grav + 2 * nb, m->dgrav, m->dgrav + nb,
                            m->curv, m->curv + nb};
    for (int k = 0; k < 7 && !rc; k++) {
      rc = h2d_field(ctx, (char*)ctx->statics + (size_t)k * sb, o, nb, src[k], ctx->ses);
      if (!rc && ctx->statics_f)
        rc = h2d_field(ctx, (char*)ctx->statics_f + (size_t)k * sizeof(float) * n, o, nb, src[k],
                       4);
    }
  }
  if (rc) return rc;
  ctx->mat_count.assign(nm, 0);
  ctx->mtab_h.assign(any_mat ? (size_t)MAX_MAT * nm : 0, MatDev{});
  if (any_mat) {  // member b's cell materials at [off_b, off_b + cells); none given: 0
    std::vector<uint8_t> mat((size_t)n, 0);
    for (int b = 0; b < nm; b++)
      if (meshes[b].mat)
        std::memcpy(mat.data() + ctx->inst_h[b].off, meshes[b].mat, (size_t)cells_of(&meshes[b]));
    HIPCHK(ctx, hipMemcpy(ctx->mat_d, mat.data(), (size_t)n, hipMemcpyHostToDevice));
  }
  if (leaves) {
    // SoA copies of the leaf geometry: x at [c], y at [n + c]
    const tg_mesh* m = &meshes[0];
    std::vector<double> cdx(2 * (size_t)n), xc(2 * (size_t)n);
    for (long long c = 0; c < n; c++) {
      cdx[c] = m->cdx[2 * c];
      cdx[n + c] = m->cdx[2 * c + 1];
      xc[c] = m->xc[2 * c];
      xc[n + c] = m->xc[2 * c + 1];
    }
    HIPCHK(ctx, hipMemcpy(ctx->leaf_nbr, nbrT.data(), sizeof(int) * nbrT.size(),
                          hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->leaf_meta, meta.data(), sizeof(uint32_t) * meta.size(),
                          hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->leaf_cdx, cdx.data(), sizeof(double) * cdx.size(),
                          hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->leaf_xc, xc.data(), sizeof(double) * xc.size(),
                          hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->leaf_cnb_off, cnb_off.data(), sizeof(int) * cnb_off.size(),
                          hipMemcpyHostToDevice));
    if (!cnb.empty())
      HIPCHK(ctx, hipMemcpy(ctx->leaf_cnb, cnb.data(), sizeof(int) * cnb.size(),
                            hipMemcpyHostToDevice));
    launch_zero(ctx->lb.kdmk, sizeof(uint32_t) * (size_t)ctx->nsuper, ctx->stream);
    ctx->leaf = LeafDev{(int)n,        ctx->leaf_nbr, ctx->leaf_meta, ctx->leaf_cdx,
                        ctx->leaf_xc,  ctx->leaf_cnb_off, ctx->leaf_cnb};
    ctx->leaves = true;
  }
  for (int v = 0; v < NSHAPE; v++)
    HIPCHK(ctx, hipMemcpyAsync(ctx->tiles_d[v], ctx->tiles_h[v].data(),
                               sizeof(TileDev) * ctx->nt[v], hipMemcpyHostToDevice, ctx->stream));
  std::vector<int> roff[NSHAPE], ridx[NSHAPE];
  for (int v = 0; v < NSHAPE; v++) {  // tiles grouped by member, row-major within it
    const std::vector<TileDev>& tl = ctx->tiles_h[v];
    ridx[v].resize(tl.size());
    for (int i = 0; i < (int)tl.size(); i++) ridx[v][i] = i;
    std::stable_sort(ridx[v].begin(), ridx[v].end(), [&](int a, int b) {
      const TileDev &p = tl[a], &q = tl[b];
      if (p.inst != q.inst) return p.inst < q.inst;
      return p.y0 != q.y0 ? p.y0 < q.y0 : p.x0 < q.x0;
    });
    roff[v].assign(nm + 1, 0);
    for (const TileDev& t : tl) roff[v][t.inst + 1]++;
    for (int b = 0; b < nm; b++) roff[v][b + 1] += roff[v][b];
    HIPCHK(ctx, hipMemcpyAsync(ctx->red_off[v], roff[v].data(), sizeof(int) * (nm + 1),
                               hipMemcpyHostToDevice, ctx->stream));
    if (!tl.empty())
      HIPCHK(ctx, hipMemcpyAsync(ctx->red_idx[v], ridx[v].data(), sizeof(int) * tl.size(),
                                 hipMemcpyHostToDevice, ctx->stream));
  }
  for (int v = 0; v < NSHAPE; v++)
    if (ctx->nb[v] > 0)
      HIPCHK(ctx, hipMemcpyAsync(ctx->blocks_d[v], ctx->blocks_h[v].data(),
                                 sizeof(BlockDev) * ctx->nb[v], hipMemcpyHostToDevice,
                                 ctx->stream));
  rc = upload_params(ctx, nullptr);
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->have_mesh = true;
  return TG_OK;
}

int reset_time(tg_ctx* ctx) {
  std::vector<double> z(ctx->ninst, 0.);
  HIPCHK(ctx, hipMemcpyAsync(ctx->time_d, z.data(), sizeof(double) * ctx->ninst,
                             hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(ctx->err_d, 0, sizeof(int), ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return TG_OK;
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
  cfg->use_graph = 0;
  cfg->store_bytes = 0;
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

int tg_material_from_angles(double phi_int, double phi_bed, tg_material* out) {
  if (!out) return TG_ERR_ARG;
  tg_params p;
  const int rc = tg_params_from_angles(phi_int, phi_bed, 1.0, 1.0, &p);
  if (rc) return rc;
  out->tan_bed = p.tan_bed;
  out->kact_root = p.kact_root;
  out->dkroot_dphi = p.dkroot_dphi;
  return TG_OK;
}

int tg_set_materials(tg_ctx* ctx, int member, int nmat, const tg_material* mats) {
  if (!ctx || nmat < 0 || nmat > TG_MAX_MATERIALS || (nmat > 0 && !mats))
    return TG_ERR_ARG;
  if (!ctx->have_mesh) return set_err(ctx, TG_ERR_STATE, "no mesh");
  if (member < 0 || member >= ctx->ninst) return set_err(ctx, TG_ERR_ARG, "no such member");
  if (!ctx->mat_d)
    return set_err(ctx, TG_ERR_STATE, "the mesh gives no per-cell materials (tg_mesh.mat)");
  for (int m = 0; m < nmat; m++)
    ctx->mtab_h[(size_t)member * MAX_MAT + m] =
        MatDev{mats[m].tan_bed, mats[m].kact_root, mats[m].dkroot_dphi, 0.};
  ctx->mat_count[member] = nmat;
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  return upload_params(ctx, nullptr);
}

int tg_set_error_grid(tg_ctx* ctx, const tg_error_grid* g) {
  if (!ctx || !g) return TG_ERR_ARG;
  if (!ctx->have_mesh) return set_err(ctx, TG_ERR_STATE, "no mesh");
  if (ctx->ninst != 1) return set_err(ctx, TG_ERR_UNSUPPORTED, "error grid: one member");
  if (!g->father || !g->kind || !g->pts || !g->pxy)
    return set_err(ctx, TG_ERR_ARG, "error grid: no father / stencil tables");
  if (g->mat && !ctx->mtab_d)
    return set_err(ctx, TG_ERR_STATE,
                   "error grid materials need a mesh with per-cell materials (tg_mesh.mat)");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  free_error_grid(ctx);
  tg_mesh m{};
  m.kind = TG_MESH_LEAVES;
  m.n = g->n;
  m.nbr = g->nbr;
  m.nproc = g->nproc;
  m.gen = g->gen;
  m.cdx = g->cdx;
  m.xc = g->xc;
  m.grav = g->grav;
  m.dgrav = g->dgrav;
  m.curv = g->curv;
  std::vector<int> nbrT;
  std::vector<uint32_t> meta;
  int rc = leaf_tables(ctx, &m, nbrT, meta);
  if (rc) return rc;
  const int n = g->n;
  const long long nf = ctx->ncell;
  for (int e = 0; e < n; e++) {
    const int k = g->kind[e];
    if (g->father[e] < 0 || g->father[e] >= nf)
      return set_err(ctx, TG_ERR_ARG, "error element " + std::to_string(e) + ": bad father");
    if (k < ERR_COPY || k > ERR_BILINEAR)
      return set_err(ctx, TG_ERR_ARG, "error element " + std::to_string(e) + ": bad rule");
    const int np = k == ERR_COPY ? 1 : (k == ERR_BARYCENTRIC ? 3 : (k == ERR_BILINEAR ? 4 : 2));
    for (int q = 0; q < np; q++)
      if (g->pts[4 * (size_t)e + q] < 0 || g->pts[4 * (size_t)e + q] >= nf)
        return set_err(ctx, TG_ERR_ARG,
                       "error element " + std::to_string(e) + ": stencil point out of range");
    if (g->mat && g->mat[e] >= MAX_MAT)
      return set_err(ctx, TG_ERR_ARG, "error element " + std::to_string(e) + ": bad material");
  }
  std::vector<double> cdx(2 * (size_t)n), xc(2 * (size_t)n), st(7 * (size_t)n);
  for (int e = 0; e < n; e++)
    for (int d = 0; d < 2; d++) {
      cdx[(size_t)d * n + e] = g->cdx[2 * (size_t)e + d];
      xc[(size_t)d * n + e] = g->xc[2 * (size_t)e + d];
    }
  for (int k = 0; k < 3; k++) std::copy(g->grav + (size_t)k * n, g->grav + (size_t)(k + 1) * n, st.begin() + (size_t)k * n);
  for (int k = 0; k < 2; k++) std::copy(g->dgrav + (size_t)k * n, g->dgrav + (size_t)(k + 1) * n, st.begin() + (size_t)(3 + k) * n);
  for (int k = 0; k < 2; k++) std::copy(g->curv + (size_t)k * n, g->curv + (size_t)(k + 1) * n, st.begin() + (size_t)(5 + k) * n);
  struct Up {
    void** d;
    const void* h;
    size_t bytes;
  } ups[] = {
      {(void**)&ctx->err_nbr, nbrT.data(), sizeof(int) * nbrT.size()},
      {(void**)&ctx->err_meta, meta.data(), sizeof(uint32_t) * meta.size()},
      {(void**)&ctx->err_cdx, cdx.data(), sizeof(double) * cdx.size()},
      {(void**)&ctx->err_xc, xc.data(), sizeof(double) * xc.size()},
      {(void**)&ctx->err_statics, st.data(), sizeof(double) * st.size()},
      {(void**)&ctx->err_mat, g->mat, g->mat ? (size_t)n : 0},
      {(void**)&ctx->err_father, g->father, sizeof(int) * (size_t)n},
      {(void**)&ctx->err_kind, g->kind, sizeof(int) * (size_t)n},
      {(void**)&ctx->err_pts, g->pts, sizeof(int) * 4 * (size_t)n},
      {(void**)&ctx->err_pxy, g->pxy, sizeof(double) * 8 * (size_t)n},
      {(void**)&ctx->err_fwd, nullptr, sizeof(double) * 9 * (size_t)nf},
      {(void**)&ctx->err_buf, nullptr, sizeof(double) * 15 * (size_t)n},
  };
  for (auto& u : ups) {
    if (!u.bytes && !u.h && u.d == (void**)&ctx->err_mat) continue;  // one material
    rc = dalloc(ctx, u.d, u.bytes);
    if (!rc && u.h) {
      const hipError_t e = hipMemcpy(*u.d, u.h, u.bytes, hipMemcpyHostToDevice);
      if (e != hipSuccess) rc = set_err(ctx, TG_ERR_HIP, hipGetErrorString(e));
    }
    if (rc) {
      free_error_grid(ctx);
      return rc;
    }
  }
  for (auto& e : ctx->err_ev)
    if (!e) HIPCHK(ctx, hipEventCreate(&e));
  ctx->err_leaf = LeafDev{n, ctx->err_nbr, ctx->err_meta, ctx->err_cdx, ctx->err_xc, nullptr,
                          nullptr};
  ctx->err_n = n;
  return TG_OK;
}

int tg_error_indicator(tg_ctx* ctx, const tg_params* prm, int step, double dt,
                       const double* prev, const double* state, const double* lambda,
                       double* residual, double* error, double* correction, double* recon) {
  if (!ctx || !prm || !prev || !state || !lambda || step < 0) return TG_ERR_ARG;
  if (!ctx->err_n) return set_err(ctx, TG_ERR_STATE, "no error grid (tg_set_error_grid)");
  if (!(dt > 0) || !std::isfinite(dt)) return set_err(ctx, TG_ERR_ARG, "dt must be > 0");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  int rc = upload_params(ctx, prm);
  if (rc) return rc;
  const long long nf = ctx->ncell;
  const size_t n = (size_t)ctx->err_n;
  const double* src[3] = {prev, state, lambda};
  for (int k = 0; k < 3; k++)
    HIPCHK(ctx, hipMemcpyAsync(ctx->err_fwd + 3 * (size_t)k * nf, src[k],
                               sizeof(double) * 3 * (size_t)nf, hipMemcpyHostToDevice,
                               ctx->stream));
  const long long ne = (long long)n;
  ErrArgs a{};
  a.n = ctx->err_n;
  a.nfwd = (int)nf;
  a.L = ctx->err_leaf;
  a.S = Statics{at<double>(ctx->err_statics, 0, ne), at<double>(ctx->err_statics, 1, ne),
                at<double>(ctx->err_statics, 2, ne), at<double>(ctx->err_statics, 3, ne),
                at<double>(ctx->err_statics, 4, ne), at<double>(ctx->err_statics, 5, ne),
                at<double>(ctx->err_statics, 6, ne)};
  a.I = ctx->inst_h[0];
  a.mat = ctx->err_mat;
  a.mtab = ctx->mtab_d;
  a.father = ctx->err_father;
  a.kind = ctx->err_kind;
  a.pts = ctx->err_pts;
  a.pxy = ctx->err_pxy;
  a.fwd = ctx->err_fwd;
  a.rec = ctx->err_buf;
  a.kact = ctx->err_buf + 9 * n;
  a.res = ctx->err_buf + 10 * n;
  a.err = ctx->err_buf + 13 * n;
  a.corr = ctx->err_buf + 14 * n;
  a.dt = dt;
  a.ft = frict_tiny(ctx, step);
  a.tiny = ctx->cfg.tiny;
  a.gamma = ctx->cfg.gamma;
  launch_error(a, ctx->stream, ctx->err_ev);
  HIPCHK(ctx, hipGetLastError());
  struct Down {
    double* h;
    const double* d;
    size_t count;
  } downs[] = {{residual, a.res, 3 * n}, {error, a.err, n}, {correction, a.corr, n},
               {recon, a.rec, 9 * n}};
  for (auto& d : downs)
    if (d.h)
      HIPCHK(ctx, hipMemcpyAsync(d.h, d.d, sizeof(double) * d.count, hipMemcpyDeviceToHost,
                                 ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < 3; k++) {
    float ms = -1.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->err_ev[k], ctx->err_ev[k + 1]));
    ctx->err_ms[k] = ms;
  }
  return TG_OK;
}

int tg_error_last_ms(const tg_ctx* ctx, double ms[3]) {
  if (!ctx || !ms) return TG_ERR_ARG;
  for (int k = 0; k < 3; k++) ms[k] = ctx->err_ms[k];
  return TG_OK;
}

int tg_create(const tg_config* cfg, tg_ctx** out) {
  if (!out) return TG_ERR_ARG;
  *out = nullptr;
  tg_ctx* ctx = new tg_ctx();
  if (cfg) ctx->cfg = *cfg;
  else tg_config_default(&ctx->cfg);
  if (ctx->cfg.checkpoint_every < 1 || ctx->cfg.max_steps < 1 ||
      (ctx->cfg.precision != TG_FP64 && ctx->cfg.precision != TG_MIXED &&
       ctx->cfg.precision != TG_FP32 && ctx->cfg.precision != TG_FWD64_ADJ32)) {
    delete ctx;
    return TG_ERR_ARG;
  }
  // tile shape: 0 x 0 keeps the built-in default variant (tg_autotune may
  // change it); an explicit shape selects its XCD-banded table for all kernels
  if (ctx->cfg.tile_x != 0 || ctx->cfg.tile_y != 0) {
    int v = -1;
    for (int k = 0; k < NSHAPE; k++)
      if (SHAPE_XCD[k] && SHAPE_TX[k] == ctx->cfg.tile_x &&
          BLOCK / SHAPE_TX[k] == ctx->cfg.tile_y)
        v = k;
    if (v < 0) {
      delete ctx;
      return TG_ERR_UNSUPPORTED;
    }
    ctx->sel[0] = ctx->sel[1] = ctx->sel[2] = v;
  } else {
    ctx->sel[0] = ctx->sel[1] = ctx->sel[2] = DEFAULT_VARIANT;
  }
  ctx->mode = ctx->cfg.precision;
  ctx->es = (ctx->mode == TG_FP32) ? 4 : 8;
  ctx->ses = fwd64(ctx) ? 8 : 4;
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
  for (auto& e : ctx->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
      tg_destroy(ctx);
      return TG_ERR_HIP;
    }
  }
  if (dalloc(ctx, (void**)&ctx->err_d, sizeof(int))) {
    tg_destroy(ctx);
    return TG_ERR_NOMEM;
  }
  if (hipMemset(ctx->err_d, 0, sizeof(int)) != hipSuccess) {
    tg_destroy(ctx);
    return TG_ERR_HIP;
  }
  *out = ctx;
  return TG_OK;
}

void tg_destroy(tg_ctx* ctx) {
  if (!ctx) return;
  // teardown: errors here have no caller to report to
  (void)hipSetDevice(ctx->cfg.device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_fields(ctx);
  dfree(ctx->err_d);
  dfree(ctx->active_d);
  for (hipEvent_t e : ctx->iev) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->seg_ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->err_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* tg_last_error(const tg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int tg_set_mesh(tg_ctx* ctx, const tg_mesh* m) {
  if (!ctx || !m) return TG_ERR_ARG;
  return set_members(ctx, 1, m, nullptr);
}

int tg_set_ensemble(tg_ctx* ctx, int nmembers, const tg_mesh* meshes, const tg_params* params) {
  if (!ctx || !meshes || !params || nmembers < 1) return TG_ERR_ARG;
  return set_members(ctx, nmembers, meshes, params);
}

int tg_num_members(const tg_ctx* ctx) { return ctx ? ctx->ninst : 0; }

int tg_member_cells(const tg_ctx* ctx, int member, int64_t* offset, int64_t* count) {
  if (!ctx || member < 0 || member >= ctx->ninst) return TG_ERR_ARG;
  const InstDev& I = ctx->inst_h[member];
  if (offset) *offset = I.off;
  if (count) *count = (int64_t)I.nx * I.ny;
  return TG_OK;
}

int tg_set_state(tg_ctx* ctx, const double* h, const double* hu, const double* hv) {
  if (!ctx || !h || !hu || !hv) return TG_ERR_ARG;
  if (!ctx->have_mesh) return set_err(ctx, TG_ERR_STATE, "no mesh");
  // a full state restores the buffers' home roles (U in the first of the
  // pair), so repeated runs from set states reuse one captured sweep
  if (ctx->U[0] != ctx->U_home)
    for (int c = 0; c < 3; c++) std::swap(ctx->U[c], ctx->Ut[c]);
  const double* src[3] = {h, hu, hv};
  for (int k = 0; k < 3; k++) {
    int rc = h2d_field(ctx, ctx->U[k], 0, ctx->ncell, src[k]);
    if (rc) return rc;
  }
  clear_marks(ctx, ctx->U[0]);
  int rc = reset_time(ctx);
  if (rc) return rc;
  ctx->step0 = 0;
  ctx->rec_steps = ctx->run_steps = 0;
  ctx->have_state = true;
  return TG_OK;
}

int tg_set_clock(tg_ctx* ctx, int step0, double time) {
  if (!ctx || step0 < 0 || !(time >= 0) || !std::isfinite(time)) return TG_ERR_ARG;
  if (!ctx->have_state) return set_err(ctx, TG_ERR_STATE, "no state");
  ctx->step0 = step0;
  std::vector<double> t(ctx->ninst, time);
  HIPCHK(ctx, hipMemcpyAsync(ctx->time_d, t.data(), sizeof(double) * ctx->ninst,
                             hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return TG_OK;
}

int tg_set_member_state(tg_ctx* ctx, int member, const double* h, const double* hu,
                        const double* hv) {
  if (!ctx || !h || !hu || !hv || member < 0 || member >= ctx->ninst) return TG_ERR_ARG;
  if (!ctx->have_mesh) return set_err(ctx, TG_ERR_STATE, "no mesh");
  const InstDev& I = ctx->inst_h[member];
  const double* src[3] = {h, hu, hv};
  for (int k = 0; k < 3; k++) {
    int rc = h2d_field(ctx, ctx->U[k], I.off, (long long)I.nx * I.ny, src[k]);
    if (rc) return rc;
  }
  clear_marks(ctx, ctx->U[0]);
  int rc = reset_time(ctx);
  if (rc) return rc;
  ctx->rec_steps = ctx->run_steps = 0;
  ctx->have_state = true;
  return TG_OK;
}

int tg_get_state(tg_ctx* ctx, double* h, double* hu, double* hv) {
  if (!ctx || !h || !hu || !hv) return TG_ERR_ARG;
  if (!ctx->have_state) return set_err(ctx, TG_ERR_STATE, "no state");
  double* dst[3] = {h, hu, hv};
  for (int k = 0; k < 3; k++) {
    int rc = d2h_field(ctx, ctx->U[k], 0, ctx->ncell, dst[k]);
    if (rc) return rc;
  }
  return TG_OK;
}

int tg_get_member_state(tg_ctx* ctx, int member, double* h, double* hu, double* hv) {
  if (!ctx || !h || !hu || !hv || member < 0 || member >= ctx->ninst) return TG_ERR_ARG;
  if (!ctx->have_state) return set_err(ctx, TG_ERR_STATE, "no state");
  const InstDev& I = ctx->inst_h[member];
  double* dst[3] = {h, hu, hv};
  for (int k = 0; k < 3; k++) {
    int rc = d2h_field(ctx, ctx->U[k], I.off, (long long)I.nx * I.ny, dst[k]);
    if (rc) return rc;
  }
  return TG_OK;
}

int tg_forward_step(tg_ctx* ctx, const tg_params* prm, tg_stats* stats) {
  if (!ctx) return TG_ERR_ARG;
  if (!ctx->have_mesh || !ctx->have_state) return set_err(ctx, TG_ERR_STATE, "mesh/state not set");
  if (ctx->rec_steps >= ctx->cfg.max_steps) return set_err(ctx, TG_ERR_ARG, "dt record full");
  int rc = upload_params(ctx, prm);
  if (rc) return rc;
  if (ctx->rec_steps == 0) {
    launch_reset_red(ctx->red, ctx->ninst, ctx->stream);
    zero_box_record(ctx);
  }
  enqueue_step(ctx, ctx->U, ctx->Ut, ctx->rec_steps, true);
  enqueue_box_max(ctx, ctx->Ut, ctx->rec_steps);
  for (int c = 0; c < 3; c++) std::swap(ctx->U[c], ctx->Ut[c]);
  HIPCHK(ctx, hipGetLastError());
  rc = check_err(ctx);
  if (rc) return rc;
  ctx->rec_steps++;
  ctx->run_steps = 0;  // single steps keep no checkpoints
  return fill_stats(ctx, stats, 1);
}

int tg_forward_run(tg_ctx* ctx, const tg_params* prm, int nsteps, tg_stats* stats) {
  if (!ctx) return TG_ERR_ARG;
  return forward_loop(ctx, prm, nsteps, nullptr, stats, stats ? ctx->ninst : 0);
}

int tg_forward_replay(tg_ctx* ctx, const tg_params* prm, int nsteps, const double* dt_seq,
                      tg_stats* stats) {
  if (!ctx || (!dt_seq && nsteps > 0)) return TG_ERR_ARG;
  return forward_loop(ctx, prm, nsteps, dt_seq, stats, stats ? ctx->ninst : 0);
}

int tg_get_dt_record(tg_ctx* ctx, double* dt_out, int capacity, int* count) {
  return tg_get_member_dt_record(ctx, 0, dt_out, capacity, count);
}

int tg_get_member_dt_record(tg_ctx* ctx, int member, double* dt_out, int capacity, int* count) {
  if (!ctx || !count || member < 0 || member >= ctx->ninst) return TG_ERR_ARG;
  *count = ctx->rec_steps;
  if (dt_out && capacity > 0) {
    int m = capacity < ctx->rec_steps ? capacity : ctx->rec_steps;
    HIPCHK(ctx, hipMemcpyAsync(dt_out, ctx->dt_rec + (long long)member * ctx->cfg.max_steps,
                               sizeof(double) * m, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  }
  return TG_OK;
}

int tg_get_trajectory_state(tg_ctx* ctx, int s, double* h, double* hu, double* hv) {
  if (!ctx || !h || !hu || !hv) return TG_ERR_ARG;
  const int N = ctx->run_steps, k = ctx->cfg.checkpoint_every, s0 = ctx->traj_s0;
  if (N <= 0) return set_err(ctx, TG_ERR_STATE, "no forward run recorded");
  if (s < 0 || s > N)
    return set_err(ctx, TG_ERR_ARG, "step outside [0, steps of the last forward run]");
  void* f[3];
  if (s == N) {
    for (int c = 0; c < 3; c++) f[c] = ctx->U[c];
  } else if (s >= s0) {
    split(ctx, traj_state(ctx, s - s0), f);
  } else {  // U^a from its checkpoint, then steps a .. s-1 with the recorded dt
    int rc = ensure_store(ctx, ctx->seg, (size_t)k);
    if (rc) return rc;
    const int a = (s / k) * k;
    void* base[3];
    split(ctx, ctx->ckpt[s / k], base);
    copy_state(ctx, ctx->seg[0], base);
    for (int t = a; t < s; t++) {
      void* in[3];
      void* out[3];
      split(ctx, ctx->seg[t - a], in);
      split(ctx, ctx->seg[t - a + 1], out);
      enqueue_step(ctx, in, out, t, false);
    }
    split(ctx, ctx->seg[s - a], f);
  }
  HIPCHK(ctx, hipGetLastError());
  int rc = check_err(ctx);
  if (rc) return rc;
  double* dst[3] = {h, hu, hv};
  for (int c = 0; c < 3; c++) {
    rc = d2h_field(ctx, f[c], 0, ctx->ncell, dst[c]);
    if (rc) return rc;
  }
  return TG_OK;
}

int tg_write_solution_files(tg_ctx* ctx, const char* dir, int rank, const uint32_t* keys,
                            const int32_t* order, int first, int last) {
  if (!ctx || !dir || !keys || rank < 0 || first < 0 || last < first) return TG_ERR_ARG;
  if (ctx->run_steps <= 0) return set_err(ctx, TG_ERR_STATE, "no forward run recorded");
  if (last > ctx->run_steps + 1)
    return set_err(ctx, TG_ERR_ARG, "last step beyond the last forward run");
  const size_t n = (size_t)ctx->ncell;
  std::vector<double> h(n), hu(n), hv(n);
  for (int s = first; s < last; s++) {
    int rc = tg_get_trajectory_state(ctx, s, h.data(), hu.data(), hv.data());
    if (rc) return rc;
    char name[64];
    std::snprintf(name, sizeof name, "/solution_%04d_%08d", rank, s);
    const std::string path = std::string(dir) + name;
    if (tg_host_write_solution(path.c_str(), (int64_t)n, keys, order, h.data(), hu.data(),
                               hv.data()) != 0)
      return set_err(ctx, TG_ERR_ARG, "cannot write " + path);
  }
  return TG_OK;
}

int tg_adjoint_run(tg_ctx* ctx, const tg_params* prm, const tg_functional* fn,
                   double* lambda0_out, double* dJ_dphibed_out) {
  if (!ctx || (fn && !fn->terminal_weight)) return TG_ERR_ARG;
  if (!fn && !ctx->have_fw) return set_err(ctx, TG_ERR_STATE, "no functional set");
  const int* steps = fn ? fn->step : (ctx->fw_steps.empty() ? nullptr : ctx->fw_steps.data());
  return adjoint_loop(ctx, prm, fn ? fn->terminal_weight : nullptr, steps, lambda0_out,
                      dJ_dphibed_out);
}

int tg_set_functional(tg_ctx* ctx, const tg_functional* fn) {
  if (!ctx || !fn || !fn->terminal_weight) return TG_ERR_ARG;
  if (!ctx->have_mesh) return set_err(ctx, TG_ERR_STATE, "no mesh");
  if (!ctx->fw) {
    int rc = dalloc(ctx, &ctx->fw, ctx->state_bytes);
    if (rc) return rc;
  }
  void* f[3];
  split(ctx, ctx->fw, f);
  for (int c = 0; c < 3; c++) {
    int rc = h2d_field(ctx, f[c], 0, ctx->ncell, fn->terminal_weight + c * ctx->ncell);
    if (rc) return rc;
  }
  // the weights' zero marks in the adjoint's tile shape (the first reverse
  // step reads them as lambda^{S}); nothing known on leaf meshes
  const int v = ctx->sel[2];
  if (!ctx->leaves && ctx->marks && ctx->nb[v] == ctx->nsuper)
    launch_copy_marking(ctx->blocks_d[v], ctx->nb[v], v, ctx->es, f, nullptr, ctx->ncell,
                        marks_of(ctx, f), nullptr, ctx->stream);
  else
    clear_marks(ctx, ctx->fw);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->have_fw = true;
  if (fn->step) ctx->fw_steps.assign(fn->step, fn->step + ctx->ninst);
  else ctx->fw_steps.clear();
  return TG_OK;
}

int tg_set_maxh_region(tg_ctx* ctx, const unsigned char* mask) {
  if (!ctx) return TG_ERR_ARG;
  if (!ctx->have_mesh) return set_err(ctx, TG_ERR_STATE, "no mesh");
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));  // earlier launches may read the lists
  dfree(ctx->box_chunks);
  dfree(ctx->box_cells);
  ctx->nbox_chunks = 0;
  ctx->box_on = false;
  ctx->box_version++;
  if (!mask) return TG_OK;
  long long nmask = 0;
  for (long long i = 0; i < ctx->ncell; i++) nmask += mask[i] != 0;
  if (nmask > 0x7fffffffLL) return set_err(ctx, TG_ERR_ARG, "box holds more than 2^31 - 1 cells");
  std::vector<long long> cells;
  cells.reserve((size_t)nmask);
  std::vector<BoxChunk> ch;
  for (int b = 0; b < ctx->ninst; b++) {
    const InstDev& I = ctx->inst_h[b];
    const long long end = I.off + (long long)I.nx * I.ny;
    for (long long i = I.off; i < end; i++) {
      if (!mask[i]) continue;
      if (ch.empty() || ch.back().inst != b || ch.back().end - ch.back().begin == BOX_CHUNK)
        ch.push_back(BoxChunk{b, (int)cells.size(), (int)cells.size(), 0});
      cells.push_back(i);
      ch.back().end++;
    }
  }
  if (!ctx->box_rec) {
    int rc = dalloc(ctx, (void**)&ctx->box_rec,
                    sizeof(double) * ctx->ninst * (size_t)ctx->cfg.max_steps);
    if (rc) return rc;
  }
  // a record made for another region is not kept
  launch_zero(ctx->box_rec, sizeof(double) * ctx->ninst * (size_t)ctx->cfg.max_steps,
              ctx->stream);
  ctx->box_on = true;
  if (ch.empty()) return TG_OK;  // no cell in the box: the record stays zero
  int rc = dalloc(ctx, (void**)&ctx->box_cells, sizeof(long long) * cells.size());
  if (rc) return rc;
  rc = dalloc(ctx, (void**)&ctx->box_chunks, sizeof(BoxChunk) * ch.size());
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpyAsync(ctx->box_cells, cells.data(), sizeof(long long) * cells.size(),
                             hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->box_chunks, ch.data(), sizeof(BoxChunk) * ch.size(),
                             hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->nbox_chunks = (int)ch.size();
  return TG_OK;
}

int tg_get_member_box_record(tg_ctx* ctx, int member, double* hmax_out, int capacity,
                             int* count) {
  if (!ctx || !count || member < 0 || member >= ctx->ninst) return TG_ERR_ARG;
  *count = ctx->box_on ? ctx->rec_steps : 0;
  if (hmax_out && capacity > 0 && *count > 0) {
    const int m = capacity < *count ? capacity : *count;
    HIPCHK(ctx, hipMemcpyAsync(hmax_out, ctx->box_rec + (long long)member * ctx->cfg.max_steps,
                               sizeof(double) * m, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  }
  return TG_OK;
}

int tg_get_maxh(tg_ctx* ctx, double* hmax, int* step) {
  if (!ctx || !hmax || !step) return TG_ERR_ARG;
  if (!ctx->box_on) return set_err(ctx, TG_ERR_STATE, "no MaxH region set");
  std::vector<double> r(ctx->rec_steps);
  for (int b = 0; b < ctx->ninst; b++) {
    int cnt = 0;
    int rc = tg_get_member_box_record(ctx, b, r.data(), (int)r.size(), &cnt);
    if (rc) return rc;
    // check_maxh: a strict '>' over the steps in order, from hmax = 0
    double m = 0.;
    int it = 0;
    for (int s = 0; s < cnt; s++)
      if (r[s] > m) {
        m = r[s];
        it = s + 1;
      }
    hmax[b] = m;
    step[b] = it;
  }
  return TG_OK;
}

int tg_get_member_dphi(tg_ctx* ctx, double* dJ_dphibed) {
  if (!ctx || !dJ_dphibed) return TG_ERR_ARG;
  if ((int)ctx->member_dphi.size() != ctx->ninst)
    return set_err(ctx, TG_ERR_STATE, "no adjoint run with the gradient");
  for (int b = 0; b < ctx->ninst; b++) dJ_dphibed[b] = ctx->member_dphi[b];
  return TG_OK;
}

int adjoint_step_impl(tg_ctx* ctx, const tg_params* prm, int s, double dt,
                      const double* lambda_next, double* lambda_out, double* dphi_member) {
  if (!ctx || !lambda_next || !lambda_out || s < 0 || s >= ctx->cfg.max_steps)
    return TG_ERR_ARG;
  if (!ctx->have_state) return set_err(ctx, TG_ERR_STATE, "no state");
  int rc = upload_params(ctx, prm);
  if (rc) return rc;
  ctx->kd_step = -1;
  const long long n = ctx->ncell;
  for (int b = 0; b < ctx->ninst; b++)  // the step's dt goes where the kernel reads it
    HIPCHK(ctx, hipMemcpyAsync(ctx->dt_rec + (long long)b * ctx->cfg.max_steps + s, &dt,
                               sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  for (int c = 0; c < 3; c++) {
    rc = h2d_field(ctx, ctx->lam[0][c], 0, n, lambda_next + c * n);
    if (rc) return rc;
  }
  clear_marks(ctx, ctx->lam[0][0]);
  const bool dphi = dphi_member != nullptr;
  if (dphi) launch_zero(ctx->dphi_part, sizeof(double) * ctx->ntiles, ctx->stream);
  enqueue_adjoint(ctx, ctx->U, ctx->lam[0], ctx->lam[1], s, dphi);
  HIPCHK(ctx, hipGetLastError());
  for (int c = 0; c < 3; c++) {
    rc = d2h_field(ctx, ctx->lam[1][c], 0, n, lambda_out + c * n);
    if (rc) return rc;
  }
  if (dphi) {
    rc = member_dphi(ctx, dphi_member);
    if (rc) return rc;
  }
  ctx->run_steps = 0;  // the dt record slot was overwritten
  return TG_OK;
}

int tg_adjoint_step(tg_ctx* ctx, const tg_params* prm, int s, double dt,
                    const double* lambda_next, double* lambda_out) {
  return adjoint_step_impl(ctx, prm, s, dt, lambda_next, lambda_out, nullptr);
}

int tg_adjoint_step_grad(tg_ctx* ctx, const tg_params* prm, int s, double dt,
                         const double* lambda_next, double* lambda_out, double* dJ_dphibed) {
  if (!dJ_dphibed) return TG_ERR_ARG;
  return adjoint_step_impl(ctx, prm, s, dt, lambda_next, lambda_out, dJ_dphibed);
}

int tg_last_timing(tg_ctx* ctx, double* forward_ms, double* adjoint_ms, double* recompute_ms) {
  if (!ctx) return TG_ERR_ARG;
  if (forward_ms) *forward_ms = ctx->fwd_ms;
  if (adjoint_ms) *adjoint_ms = ctx->adj_ms;
  if (recompute_ms) *recompute_ms = ctx->rec_ms;
  return TG_OK;
}

int tg_last_stored_sweep(tg_ctx* ctx, double* stored_ms, int* stored_steps) {
  if (!ctx) return TG_ERR_ARG;
  if (stored_ms) *stored_ms = ctx->stored_ms;
  if (stored_steps) *stored_steps = ctx->stored_steps;
  return TG_OK;
}

// Average launch time of kernel `which` with tile variant v (-1: the
// selected one), over reps back-to-back launches on the current state.
int time_kernel_variant(tg_ctx* ctx, int which, int v, int reps, double* avg_ms) {
  const int keep[3] = {ctx->sel[0], ctx->sel[1], ctx->sel[2]};
  ctx->kd_step = -1;
  ctx->bytes_off = true;
  if (v >= 0) ctx->sel[which] = v;
  // every variant reads U and writes scratch (Ut or lam[1]) with step 0's
  // dt; the adjoint runs with its bed-friction gradient term, as in
  // tg_adjoint_run with dJ_dphibed requested (its partial sums are scratch
  // here and cleared by the next tg_adjoint_run)
  auto one = [&](int w) {
    if (w == 2) {
      enqueue_adjoint(ctx, ctx->U, ctx->lam[0], ctx->lam[1], 0, true);
    } else if (fwd64(ctx)) {
      StepArgs a{};
      const long long n = ctx->ncell;
      a.inst = ctx->inst_d;
      use_tiles(ctx, w, a);
      a.U = Fields{(double*)ctx->U[0], (double*)ctx->U[1], (double*)ctx->U[2]};
      a.S = Statics{at<double>(ctx->statics, 0, n), at<double>(ctx->statics, 1, n),
                    at<double>(ctx->statics, 2, n), at<double>(ctx->statics, 3, n),
                    at<double>(ctx->statics, 4, n), at<double>(ctx->statics, 5, n),
                    at<double>(ctx->statics, 6, n)};
      a.kact = (double*)ctx->kact; a.dt_rec = ctx->dt_rec; a.max_steps = ctx->cfg.max_steps;
      a.ft = frict_tiny(ctx, 0); a.tiny = ctx->cfg.tiny; a.gamma = ctx->cfg.gamma;
      a.h_out = (double*)ctx->Ut[0]; a.hu_out = (double*)ctx->Ut[1]; a.hv_out = (double*)ctx->Ut[2];
      a.err = ctx->err_d;
      a.tile_list = ctx->act_list;
      a.tile_count = ctx->act_count;
      a.mk_in = marks_of(ctx, ctx->U);
      a.mk_out = marks_of(ctx, ctx->Ut);
      a.zbytes = 0;  // (timing: see bytes_off)
      a.active = inst_counter(ctx, w, true);
      a.short_cells = inst_short(ctx, w, true);
      inst_mark(ctx, w, true);
      if (ctx->leaves) {
        if (w == 0) launch_leaf_coef(a, ctx->leaf, false, ctx->stream);
        else launch_leaf_update(a, ctx->leaf, ctx->stream);
      } else if (w == 0) {
        launch_coef(a, false, ctx->stream);
      } else {
        launch_update(a, ctx->stream);
      }
      inst_mark(ctx, w, false);
    } else if (ctx->es == 8) {
      StepArgsM<double> a = step_args_m<double>(ctx, ctx->U, ctx->Ut, 0);
      use_tiles(ctx, w, a);
      if (w == 0) launch_coef_m<double>(a, false, ctx->stream);
      else launch_update_m<double>(a, ctx->stream);
    } else {
      StepArgsM<float> a = step_args_m<float>(ctx, ctx->U, ctx->Ut, 0);
      use_tiles(ctx, w, a);
      if (w == 0) launch_coef_m<float>(a, false, ctx->stream);
      else launch_update_m<float>(a, ctx->stream);
    }
  };
  one(0);  // kact of the current state
  one(which);  // untimed first launch of this variant
  int rc = TG_OK;
  float ms = 0;
  if (hipEventRecord(ctx->ev[0], ctx->stream) != hipSuccess) rc = TG_ERR_HIP;
  for (int r = 0; r < reps && !rc; r++) one(which);
  if (!rc && (hipEventRecord(ctx->ev[1], ctx->stream) != hipSuccess ||
              hipGetLastError() != hipSuccess || hipEventSynchronize(ctx->ev[1]) != hipSuccess ||
              hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) != hipSuccess))
    rc = set_err(ctx, TG_ERR_HIP, "kernel timing failed");
  for (int k = 0; k < 3; k++) ctx->sel[k] = keep[k];
  ctx->bytes_off = false;
  // the scratch buffers carry marks in the timed variant's shape
  if (ctx->marks) {
    clear_marks(ctx, ctx->Ut[0]);
    clear_marks(ctx, ctx->lam[1][0]);
  }
  if (rc) return rc;
  *avg_ms = ms / reps;
  return check_err(ctx);
}

int tg_time_kernel(tg_ctx* ctx, const tg_params* prm, int which, int reps, double* avg_ms) {
  if (!ctx || !avg_ms || reps < 1 || which < 0 || which > 2) return TG_ERR_ARG;
  if (!ctx->have_mesh || !ctx->have_state || ctx->rec_steps < 1)
    return set_err(ctx, TG_ERR_STATE, "needs a completed forward run");
  int rc = upload_params(ctx, prm);
  if (rc) return rc;
  return time_kernel_variant(ctx, which, -1, reps, avg_ms);
}

int tg_autotune(tg_ctx* ctx, const tg_params* prm, int reps, int* variant_out, double* ms_out) {
  if (!ctx || reps < 1) return TG_ERR_ARG;
  if (!ctx->have_mesh || !ctx->have_state || ctx->rec_steps < 1)
    return set_err(ctx, TG_ERR_STATE, "needs a completed forward run");
  int rc = upload_params(ctx, prm);
  if (rc) return rc;
  if (ctx->leaves) {  // one thread per leaf: no tile variants to choose from
    for (int w = 0; w < 3; w++) {
      ctx->sel[w] = 0;
      if (variant_out) variant_out[w] = 0;
      if (ms_out)
        for (int v = 0; v < NSHAPE; v++) ms_out[w * NSHAPE + v] = 0.;
    }
    return TG_OK;
  }
  // the adjoint variant reads lam[0]: give it defined values (zeros)
  const size_t fb = (size_t)ctx->es * ctx->ncell;
  for (int c = 0; c < 3; c++) HIPCHK(ctx, hipMemsetAsync(ctx->lam[0][c], 0, fb, ctx->stream));
  clear_marks(ctx, ctx->lam[0][0]);
  for (int w = 0; w < 3; w++) {
    double best = HUGE_VAL;
    int bv = ctx->sel[w];
    for (int v = 0; v < NSHAPE; v++) {
      double ms = 0;
      rc = time_kernel_variant(ctx, w, v, reps, &ms);
      if (rc) return rc;
      if (ms_out) ms_out[w * NSHAPE + v] = ms;
      if (ms < best) { best = ms; bv = v; }
    }
    ctx->sel[w] = bv;
    if (variant_out) variant_out[w] = bv;
  }
  return TG_OK;
}

int tg_device_memory(int device, int64_t* free_bytes, int64_t* total_bytes) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return TG_ERR_NODEVICE;
  size_t f = 0, t = 0;
  if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&f, &t) != hipSuccess)
    return TG_ERR_HIP;
  if (free_bytes) *free_bytes = (int64_t)f;
  if (total_bytes) *total_bytes = (int64_t)t;
  return TG_OK;
}

int tg_tile_variant(int v, int* tile_x, int* tile_y, int* xcd_banded) {
  if (v < 0 || v >= NSHAPE) return TG_ERR_ARG;
  if (tile_x) *tile_x = SHAPE_TX[v];
  if (tile_y) *tile_y = BLOCK / SHAPE_TX[v];
  if (xcd_banded) *xcd_banded = SHAPE_XCD[v] ? 1 : 0;
  return TG_OK;
}

int tg_set_tile_selection(tg_ctx* ctx, const int* variant) {
  if (!ctx || !variant) return TG_ERR_ARG;
  for (int k = 0; k < 3; k++)
    if (variant[k] < 0 || variant[k] >= NSHAPE) return TG_ERR_ARG;
  for (int k = 0; k < 3; k++) ctx->sel[k] = variant[k];
  return TG_OK;
}

int tg_get_tile_selection(const tg_ctx* ctx, int* variant_out) {
  if (!ctx || !variant_out) return TG_ERR_ARG;
  for (int k = 0; k < 3; k++) variant_out[k] = ctx->sel[k];
  return TG_OK;
}

int64_t tg_num_cells(const tg_ctx* ctx) { return ctx ? (int64_t)ctx->ncell : 0; }

int tg_set_instrumentation(tg_ctx* ctx, int on) {
  if (!ctx) return TG_ERR_ARG;
  int rc = inst_flush(ctx);
  if (rc) return rc;
  if (on && !ctx->active_d) {
    rc = dalloc(ctx, (void**)&ctx->active_d, 6 * sizeof(unsigned long long));
    if (rc) return rc;
  }
  if (on) {
    HIPCHK(ctx, hipMemsetAsync(ctx->active_d, 0, 6 * sizeof(unsigned long long), ctx->stream));
    for (int k = 0; k < 3; k++) {
      ctx->inst_ms[k] = 0;
      ctx->inst_launches[k] = 0;
      ctx->inst_counted[k] = true;
      ctx->inst_short_counted[k] = true;
    }
  }
  ctx->instrument = on != 0;
  return TG_OK;
}

int tg_get_instrumentation(tg_ctx* ctx, double* kernel_ms, int64_t* launches,
                           int64_t* full_cells, int64_t* short_cells) {
  if (!ctx) return TG_ERR_ARG;
  if (!ctx->active_d) return set_err(ctx, TG_ERR_STATE, "instrumentation was never on");
  int rc = inst_flush(ctx);
  if (rc) return rc;
  unsigned long long cnt[6] = {0, 0, 0, 0, 0, 0};
  HIPCHK(ctx, hipMemcpyAsync(cnt, ctx->active_d, sizeof(cnt), hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < 3; k++) {
    if (kernel_ms) kernel_ms[k] = ctx->inst_ms[k];
    if (launches) launches[k] = ctx->inst_launches[k];
    if (full_cells) full_cells[k] = ctx->inst_counted[k] ? (int64_t)cnt[k] : -1;
    if (short_cells) short_cells[k] = ctx->inst_short_counted[k] ? (int64_t)cnt[3 + k] : -1;
  }
  return TG_OK;
}

}  // extern "C"
