// tg_step_noalloc: one order-1 forward step of a uniform block with every
// intermediate kept in a fixed device array of this file (no device memory
// is allocated or freed by the call).  It runs the kernels of
// tg_forward_step -- coefficient / CFL pass, dt, classifying pass and
// update -- on caller-owned device arrays, so its results are those of
// tg_forward_step on the same state, bit for bit.
//
// Scratch: the member record, the tile and mark-block tables of the 64x4
// row-major tile variant, the CFL reduction cells, the dt / time / error
// words, the update's tile list and kact of every cell.  The array is
// per device and per process; calls on one device share it, so calls that
// overlap on one device race (the call is synchronous, and a process-wide
// mutex serialises the host side).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "../../include/tg.h"
#include "tg_kernels.h"

using namespace tg;

namespace {

constexpr long long NA_MAX_CELLS = TG_NOALLOC_MAX_CELLS;
constexpr int NA_MAX_TILES = TG_NOALLOC_MAX_CELLS / 256 * 2;  // 64x4 tiles, partial edges
constexpr int NA_MAX_BLOCKS = NA_MAX_TILES / 2;                // 64x16 mark blocks
constexpr int NA_MAX_DEVICES = 64;

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// byte offsets into the scratch array
constexpr size_t OFF_INST = 0;
constexpr size_t OFF_RED = OFF_INST + align256(sizeof(InstDev));
constexpr size_t OFF_MISC = OFF_RED + align256(sizeof(RedDev));  // dt, time, err, count
constexpr size_t OFF_TILES = OFF_MISC + 256;
constexpr size_t OFF_BLOCKS = OFF_TILES + align256(sizeof(TileDev) * NA_MAX_TILES);
constexpr size_t OFF_LIST = OFF_BLOCKS + align256(sizeof(BlockDev) * NA_MAX_BLOCKS);
constexpr size_t OFF_KACT = OFF_LIST + align256(sizeof(int) * NA_MAX_TILES);
constexpr size_t NA_BYTES = OFF_KACT + align256(sizeof(double) * NA_MAX_CELLS);

__device__ __align__(256) unsigned char scratch[NA_BYTES];

std::mutex g_mu;
unsigned char* g_ptr[NA_MAX_DEVICES] = {};  // scratch's address on each device

// host staging of the tables (static storage: nothing allocated per call)
InstDev g_inst;
TileDev g_tiles[NA_MAX_TILES];
BlockDev g_blocks[NA_MAX_BLOCKS];
struct Misc {
  double dt_rec, time;
  int err, tile_count;
};

int device_scratch(int device, unsigned char** out) {
  if (device < 0 || device >= NA_MAX_DEVICES) return TG_ERR_ARG;
  if (!g_ptr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TG_ERR_HIP;
    g_ptr[device] = (unsigned char*)p;
  }
  *out = g_ptr[device];
  return TG_OK;
}

// The tile / block tables of the 64x4 row-major variant for one nx x ny
// member (as tg_set_mesh builds them for variant 0).
void build_tables(int nx, int ny, int* ntiles, int* nblocks) {
  const int TX = SHAPE_TX[0], TY = BLOCK / TX;
  const int nsx = (nx + MARK_W - 1) / MARK_W, nsy = (ny + MARK_H - 1) / MARK_H;
  int nt = 0;
  for (int ty = 0; ty < ny; ty += TY)
    for (int tx = 0; tx < nx; tx += TX) {
      const int bx = tx / MARK_W, by = ty / MARK_H, mw = by * nsx + bx;
      g_tiles[nt++] = TileDev{0, tx, ty, mw, bx > 0 ? mw - 1 : -1, bx + 1 < nsx ? mw + 1 : -1,
                              by > 0 ? mw - nsx : -1, by + 1 < nsy ? mw + nsx : -1};
    }
  const int nb = nsx * nsy;
  for (int by = 0; by < nsy; by++)
    for (int bx = 0; bx < nsx; bx++) {
      const int mw = by * nsx + bx;
      g_blocks[mw] = BlockDev{0, mw, bx > 0 ? mw - 1 : -1, bx + 1 < nsx ? mw + 1 : -1,
                              by > 0 ? mw - nsx : -1, by + 1 < nsy ? mw + nsx : -1,
                              {-1, -1, -1, -1}, bx * MARK_W, by * MARK_H, nx, ny, 0};
    }
  for (int i = 0; i < nt; i++) {
    const TileDev& t = g_tiles[i];
    const int byte = ((t.y0 % MARK_H) / TY) * (MARK_W / TX) + (t.x0 % MARK_W) / TX;
    g_blocks[t.mw].tile[byte] = i;
  }
  *ntiles = nt;
  *nblocks = nb;
}

int step_on_device(const tg_config* cfg, const tg_noalloc_block* blk, const tg_params* prm,
                   int step, double time, const double* d_h, const double* d_hu,
                   const double* d_hv, double* d_h_out, double* d_hu_out, double* d_hv_out,
                   hipStream_t s, double* dt_out);

#define NA_CHK(call)                                  \
  do {                                                \
    if ((call) != hipSuccess) return TG_ERR_HIP;      \
  } while (0)

}  // namespace

extern "C" int tg_step_noalloc(const tg_config* cfg, const tg_noalloc_block* blk,
                               const tg_params* prm, int step, double time,
                               const double* d_h, const double* d_hu, const double* d_hv,
                               double* d_h_out, double* d_hu_out, double* d_hv_out,
                               void* stream, double* dt_out) {
  if (!cfg || !blk || !prm || !d_h || !d_hu || !d_hv || !d_h_out || !d_hu_out || !d_hv_out ||
      !blk->statics || step < 0)
    return TG_ERR_ARG;
  if (blk->nx < 1 || blk->ny < 1 || !(blk->dx > 0) || !(blk->dy > 0)) return TG_ERR_ARG;
  const long long n = (long long)blk->nx * blk->ny;
  const long long ntiles_need = (long long)((blk->nx + 63) / 64) * ((blk->ny + 3) / 4);
  if (n > NA_MAX_CELLS || ntiles_need > NA_MAX_TILES) return TG_ERR_TOO_LARGE;
  std::lock_guard<std::mutex> lock(g_mu);
  int dev = -1;
  NA_CHK(hipGetDevice(&dev));
  if (cfg->device != dev) NA_CHK(hipSetDevice(cfg->device));
  const int rc = step_on_device(cfg, blk, prm, step, time, d_h, d_hu, d_hv, d_h_out, d_hu_out,
                                d_hv_out, (hipStream_t)stream, dt_out);
  if (cfg->device != dev) (void)hipSetDevice(dev);  // the caller's device again
  return rc;
}

extern "C" long long tg_noalloc_scratch_bytes(void) { return (long long)NA_BYTES; }

namespace {

int step_on_device(const tg_config* cfg, const tg_noalloc_block* blk, const tg_params* prm,
                   int step, double time, const double* d_h, const double* d_hu,
                   const double* d_hv, double* d_h_out, double* d_hu_out, double* d_hv_out,
                   hipStream_t s, double* dt_out) {
  const long long n = (long long)blk->nx * blk->ny;
  unsigned char* base = nullptr;
  int rc = device_scratch(cfg->device, &base);
  if (rc) return rc;

  int ntiles = 0, nblocks = 0;
  build_tables(blk->nx, blk->ny, &ntiles, &nblocks);
  std::memset(&g_inst, 0, sizeof(g_inst));
  g_inst.nx = blk->nx; g_inst.ny = blk->ny; g_inst.off = 0;
  g_inst.nsx = (blk->nx + MARK_W - 1) / MARK_W;
  g_inst.nsy = (blk->ny + MARK_H - 1) / MARK_H;
  g_inst.soff = 0;
  g_inst.dx = blk->dx; g_inst.dy = blk->dy;
  g_inst.inv_dx = 1 / blk->dx; g_inst.inv_sx = 1 / (blk->dx + blk->dx);
  g_inst.inv_dy = 1 / blk->dy; g_inst.inv_sy = 1 / (blk->dy + blk->dy);
  g_inst.sin_int = prm->sin_int; g_inst.tan_bed = prm->tan_bed;
  g_inst.kc1 = prm->kact_c1; g_inst.kroot = prm->kact_root; g_inst.dkroot = prm->dkroot_dphi;
  g_inst.eps = prm->epsilon; g_inst.gscale = prm->gravity_scale;
  g_inst.maxtime = prm->maxtime; g_inst.nextoutput = prm->nextoutput;
  g_inst.nextsave = prm->nextsave;
  const Misc misc{0.0, time, 0, 0};

  NA_CHK(hipMemcpyAsync(base + OFF_INST, &g_inst, sizeof(InstDev), hipMemcpyHostToDevice, s));
  NA_CHK(hipMemcpyAsync(base + OFF_MISC, &misc, sizeof(Misc), hipMemcpyHostToDevice, s));
  NA_CHK(hipMemcpyAsync(base + OFF_TILES, g_tiles, sizeof(TileDev) * ntiles,
                        hipMemcpyHostToDevice, s));
  NA_CHK(hipMemcpyAsync(base + OFF_BLOCKS, g_blocks, sizeof(BlockDev) * nblocks,
                        hipMemcpyHostToDevice, s));
  Misc* md = (Misc*)(base + OFF_MISC);

  StepArgs a{};
  a.inst = (const InstDev*)(base + OFF_INST);
  a.tiles = (const TileDev*)(base + OFF_TILES);
  a.ntiles = ntiles;
  a.shape = 0;
  a.blocks = (const BlockDev*)(base + OFF_BLOCKS);
  a.nblocks = nblocks;
  a.U = Fields{d_h, d_hu, d_hv};
  const double* st = blk->statics;
  a.S = Statics{st, st + n, st + 2 * n, st + 3 * n, st + 4 * n, st + 5 * n, st + 6 * n};
  a.kact = (double*)(base + OFF_KACT);
  a.red = (RedDev*)(base + OFF_RED);
  a.dt_rec = &md->dt_rec;
  a.time = &md->time;
  a.max_steps = 1;
  a.step = 0;
  a.ft = (step < cfg->frict_tiny_switch) ? cfg->frict_tiny_early : cfg->frict_tiny_late;
  a.tiny = cfg->tiny; a.gamma = cfg->gamma;
  a.cfl_inner = cfg->cfl_inner; a.cfl_outer = cfg->cfl_outer;
  a.h_out = d_h_out; a.hu_out = d_hu_out; a.hv_out = d_hv_out;
  a.err = &md->err;
  a.active = nullptr;
  a.tile_list = (int*)(base + OFF_LIST);
  a.tile_count = &md->tile_count;
  a.dt_clears_count = 1;
  a.mk_in = nullptr;   // no zero marks: every tile is classified from its cells
  a.mk_out = nullptr;
  a.short_cells = nullptr;

  launch_reset_red(a.red, 1, s);
  launch_coef(a, true, s);
  launch_dt(a, 1, s);
  launch_update(a, s);
  NA_CHK(hipGetLastError());
  Misc back{};
  NA_CHK(hipMemcpyAsync(&back, md, sizeof(Misc), hipMemcpyDeviceToHost, s));
  NA_CHK(hipStreamSynchronize(s));
  if (back.err & 1) return TG_ERR_NONFINITE;
  if (back.err & 2) return TG_ERR_EIGEN;
  if (dt_out) *dt_out = back.dt_rec;
  return TG_OK;
}

}  // namespace
