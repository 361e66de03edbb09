// FP64 kernels of the forward step and the per-step discrete adjoint on
// uniform blocks of cells.  One workgroup = one 64x4 tile of one instance
// (tile table), so an ensemble of uneven blocks runs in one launch.
//
// Forward step s (replaces step.C:38-232, order 1):
//   k_coef   : forward slopes -> kact (getcoef.f) + per-instance CFL min /
//              hmax reduction (get_coef_and_eigen.C:29-180)
//   k_dt     : dt = 0.5*min(0.5*min(dx/ev), 0.9*sqrt(hmax eps Gs/9.8)), incrtime
//   k_update : slopes, HLL edge fluxes with outflow copies, update_states
// Reverse step s (replaces calc_jacobian + calc_adjoint, dual_solver.C:200-206):
//   k_adjoint: gathers lambda^s_c = g_c - sum_j J_j[slot->c]^T lambda^{s+1}_j,
//              recomputing every Jacobian block from U^s in registers.
#include "tg_kernels.h"
#include "tg_math.h"

namespace tg {

namespace {

__device__ __forceinline__ unsigned long long dbits(double v) {
  return (unsigned long long)__double_as_longlong(v);
}
__device__ __forceinline__ double bitsd(unsigned long long v) {
  return __longlong_as_double((long long)v);
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o, 64);
    v = (w < v) ? w : v;
  }
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o, 64);
    v = (w > v) ? w : v;
  }
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

struct Cell {
  int i, j;
  long long c;
  bool act;
  bool hasE, hasW, hasN, hasS;
};

template <int TX>
__device__ __forceinline__ Cell cell_of(const TileDev& t, const InstDev& I) {
  Cell q;
  q.i = t.x0 + (int)(threadIdx.x % TX);
  q.j = t.y0 + (int)(threadIdx.x / TX);
  q.act = q.i < I.nx && q.j < I.ny;
  q.c = I.off + (long long)q.j * I.nx + q.i;
  q.hasE = q.i + 1 < I.nx;
  q.hasW = q.i > 0;
  q.hasN = q.j + 1 < I.ny;
  q.hasS = q.j > 0;
  return q;
}

// ------------------------------------------------------------- zero marks --

__device__ __forceinline__ bool bitzero3(double a, double b, double c) {
  return (dbits(a) | dbits(b) | dbits(c)) == 0ull;
}
// Records in the output marks whether the tile holds +0.0 only (all: the
// workgroup's reduction of "this cell is +0.0 in all three fields or
// outside the member").  Thread 0 stores the tile's byte.
template <int TX>
__device__ __forceinline__ void mark_store(uint32_t* mk, const TileDev& t, bool all) {
  if (threadIdx.x == 0) ((unsigned char*)(mk + t.mw))[mark_byte<TX>(t)] = all ? 1 : 0;
}

// Skips decided from the marks alone, for a whole mark block.
// Coefficient pass: the block of U is +0.0 -- h = +0.0 < tiny: no kact
// written, nothing added to the CFL reduction.
__device__ __forceinline__ bool coef_skip(const uint32_t* mk, int mw) {
  return mk[mw] == MARK_FULL;
}
// Update: the block and the four side blocks of U are +0.0 -- every cell
// and side neighbour has h < tiny, so the fluxes are zero and there are no
// sources, and U^{s+1} = U^s - dtdx*(0 - 0) - dtdy*(0 - 0) = +0.0 for a
// finite dt -- and the block of the output already holds +0.0.
__device__ __forceinline__ bool update_skip(const uint32_t* in, const uint32_t* out,
                                            const BlockDev& B, double dt) {
  return in[B.mw] == MARK_FULL && out[B.mw] == MARK_FULL &&
         (B.mwW < 0 || in[B.mwW] == MARK_FULL) && (B.mwE < 0 || in[B.mwE] == MARK_FULL) &&
         (B.mwS < 0 || in[B.mwS] == MARK_FULL) && (B.mwN < 0 || in[B.mwN] == MARK_FULL) &&
         isfinite(dt);
}
// Adjoint: the block of U^s is +0.0 (void cells: lambda^s = +0.0), that of
// lambda^{s+1} too when the gradient is wanted (no bed-friction terms), and
// the block of lambda^s already holds +0.0.
__device__ __forceinline__ bool adjoint_skip(const uint32_t* mu, const uint32_t* ml,
                                             const uint32_t* out, int mw) {
  return mu[mw] == MARK_FULL && out[mw] == MARK_FULL && (!ml || ml[mw] == MARK_FULL);
}
// the byte of block word mw that tile byte b owns
__device__ __forceinline__ void mark_byte_store(uint32_t* mk, int mw, int b, bool zero) {
  ((unsigned char*)(mk + mw))[b] = zero ? 1 : 0;
}
// The cells of a block's four tiles this thread handles (one per tile).
template <int TX>
__device__ __forceinline__ void block_cells(const BlockDev& B, Cell q[4]) {
  constexpr int TY = BLOCK / TX, PER_ROW = MARK_W / TX;
  InstDev I;  // the fields cell_of reads, carried by the block record
  I.nx = B.nx; I.ny = B.ny; I.off = B.off;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    q[b].act = false;
    if (B.tile[b] < 0) continue;
    // tile byte b (mark_byte) from the block origin: no tile record to load
    TileDev t{};
    t.inst = B.inst;
    t.x0 = B.x0 + (b % PER_ROW) * TX;
    t.y0 = B.y0 + (b / PER_ROW) * TY;
    q[b] = cell_of<TX>(t, I);
  }
}

// ---------------------------------------------------------------- forward --

// Limited slope from its two one-sided differences dp = (u+ - u)/dx and
// dm = (u - u-)/dx (Element::get_slopes, element2.C:886-897; same operands
// as slope_fwd).
__device__ __forceinline__ double slope_from(double dp, double dm, double dx, double gamma,
                                             double tiny) {
  double dc = (dp * dx + dm * dx) / (dx + dx);
  return .5 * (c_sgn(dp, tiny) + c_sgn(dm, tiny)) *
         dmin3(gamma * dabs(dp), gamma * dabs(dm), dabs(dc));
}


// Coefficient pass, one workgroup per mark block (one cell of each of its
// four tiles per thread): forward slopes -> kact (getcoef.f) and the
// per-instance CFL min / hmax reduction (get_coef_and_eigen.C:29-180).
template <int TX>
__global__ __launch_bounds__(BLOCK) void k_coef(StepArgs a, int reduce) {
  // (the block table is indexed by mark word)
  if (a.mk_in && coef_skip(a.mk_in, (int)blockIdx.x)) return;
  const BlockDev B = a.blocks[blockIdx.x];
  const InstDev I = a.inst[B.inst];
  Cell q[4];
  block_cells<TX>(B, q);
  // a tile the marks show to hold +0.0 has h < tiny everywhere: nothing to
  // read or write
  const uint32_t zin = (a.mk_in && a.zbytes) ? a.mk_in[blockIdx.x] : 0u;
#pragma unroll
  for (int b = 0; b < 4; b++)
    if ((zin >> (8 * b)) & 1u) q[b].act = false;
  double h[4];
#pragma unroll
  for (int b = 0; b < 4; b++) h[b] = q[b].act ? a.U.h[q[b].c] : 0.;
  double minev = HUGE_VAL, hmax = 0.;
  int nwet = 0, nshort = 0;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    if (!q[b].act) continue;
    const long long c = q[b].c, nx = I.nx;
    if (h[b] > a.tiny) {
      nwet++;
      const double hc = h[b], hu = a.U.hu[c], hv = a.U.hv[c];
      double hx = 0., hux = 0., hy = 0., hvy = 0.;
      if (q[b].hasE && q[b].hasW && q[b].hasN && q[b].hasS) {  // boundary cells: zero slopes
        hx = slope_fwd(a.U.h[c + 1], hc, a.U.h[c - 1], I.dx, a.gamma, a.tiny);
        hux = slope_fwd(a.U.hu[c + 1], hu, a.U.hu[c - 1], I.dx, a.gamma, a.tiny);
        hy = slope_fwd(a.U.h[c + nx], hc, a.U.h[c - nx], I.dy, a.gamma, a.tiny);
        hvy = slope_fwd(a.U.hv[c + nx], hv, a.U.hv[c - nx], I.dy, a.gamma, a.tiny);
      }
      const double kroot = mat_of(I, a.mat, a.mtab, c).kroot;
      const double k = getcoef(hc, hu, hv, hx, hux, hy, hvy, I.kc1, kroot, I.eps, a.tiny);
      if (reduce) {
        double ev = eigen(hc, hu, hv, k, a.S.gz[c], a.tiny);
        if (ev > 1000000000.) atomicOr(a.err, 2);
        double mdx = (I.dx > I.dy) ? I.dy : I.dx;
        const double m = mdx / ev;
        minev = (m < minev) ? m : minev;
        hmax = (hc > hmax) ? hc : hmax;
      }
      a.kact[c] = k;
    } else {
      nshort++;
      // kact is read only where h >= tiny (the physical flux is zero below
      // tiny, update_states uses kact only above it): at or below tiny the
      // value is epsilon, and below tiny it is not written
      if (!(h[b] < a.tiny)) a.kact[c] = I.eps;
    }
  }
  if (a.active) {
    const int cnt = (int)wave_sum((double)nwet);
    if (threadIdx.x % 64 == 0 && cnt) atomicAdd(a.active, (unsigned long long)cnt);
  }
  if (a.short_cells) {
    const int cnt = (int)wave_sum((double)nshort);
    if (threadIdx.x % 64 == 0 && cnt) atomicAdd(a.short_cells, (unsigned long long)cnt);
  }
  if (reduce && __syncthreads_or(nwet > 0)) {
    __shared__ double smin[BLOCK / 64], smax[BLOCK / 64];
    minev = wave_min(minev);
    hmax = wave_max(hmax);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) { smin[w] = minev; smax[w] = hmax; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < BLOCK / 64; k++) {
        minev = (smin[k] < minev) ? smin[k] : minev;
        hmax = (smax[k] > hmax) ? smax[k] : hmax;
      }
      if (minev < HUGE_VAL) atomicMin(&a.red[B.inst].min_ev, dbits(minev));
      if (hmax > 0.) atomicMax(&a.red[B.inst].hmax, dbits(hmax));
    }
  }
}

__global__ void k_reset_red(RedDev* red, int ninst) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < ninst) {
    red[b].min_ev = dbits(10000000.);  // get_coef_and_eigen.C:35
    red[b].hmax = dbits(0.);
  }
}

// Plain fill / copy kernels: used instead of memset / memcpy nodes so a
// captured sweep holds kernel nodes only.
__global__ void k_zero_u32(uint32_t* p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0u;
}
__global__ void k_copy_u32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                           long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// MaxH box record (check_maxh, calc_adjoint.C:443-465): each workgroup
// takes one chunk of a member's box cells and raises rec[member][step] to
// the largest h > 0 among them.  The running value starts at +0.0 and only
// grows by a strict '>' (NaN never wins), so it is a non-negative double
// whose bit pattern orders like the integer: an integer atomicMax is exact.
template <class S>
__global__ void __launch_bounds__(256) k_box_max(const BoxChunk* __restrict__ ch,
                                                 const long long* __restrict__ cells,
                                                 const S* __restrict__ h,
                                                 unsigned long long* rec, int max_steps,
                                                 int step) {
  __shared__ double wm[4];
  const BoxChunk c = ch[blockIdx.x];
  double m = 0.;
  for (int i = c.begin + (int)threadIdx.x; i < c.end; i += 256) {
    const double v = (double)h[cells[i]];
    m = (v > m) ? v : m;
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; w++) m = (wm[w] > m) ? wm[w] : m;
    if (m > 0.) atomicMax(rec + (long long)c.inst * max_steps + step, dbits(m));
  }
}

// The three field ranges [off, off + count) of one member (words of `es`
// bytes in buffers of `field_bytes` per field): copied from src, or zeroed
// when src is null; the member's zero-mark words [mw0, mw0 + nmw) of dst are
// cleared (nothing known) after a copy.
__global__ void k_member_fill(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                              long long field_words, long long off_words, long long nwords,
                              uint32_t* mk, long long mw0, long long nmw) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  for (int f = 0; f < 3; f++) {
    const long long base = f * field_words + off_words;
    for (long long i = t0; i < nwords; i += stride)
      dst[base + i] = src ? src[base + i] : 0u;
  }
  if (src && mk)
    for (long long i = t0; i < nmw; i += stride) mk[mw0 + i] = 0u;
}

__global__ void k_dt(StepArgs a, int ninst) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b == 0 && a.dt_clears_count) *a.tile_count = 0;  // the update's list starts empty
  if (b >= ninst) return;
  const InstDev I = a.inst[b];
  double min_ev = bitsd(a.red[b].min_ev), hmax = bitsd(a.red[b].hmax);
  double dt0 = a.cfl_inner * min_ev;
  double dt1 = -0.9 * sqrt(hmax * I.eps * I.gscale / 9.8);
  double m = (dt0 > -dt1) ? -dt1 : dt0;
  double dt = a.cfl_outer * m;
  if (dt == 0.0) dt = 0.5 * HUGE_VAL;
  // TimeProps::incrtime, properties.h:397-410
  double time = a.time[b];
  if (time + dt > I.nextoutput) dt = I.nextoutput - time;
  if (time + dt > I.nextsave) dt = I.nextsave - time;
  if (time + dt > I.maxtime) dt = I.maxtime - time;
  a.time[b] = time + dt;
  a.dt_rec[(long long)b * a.max_steps + a.step] = dt;
  a.red[b].min_ev = dbits(10000000.);
  a.red[b].hmax = dbits(0.);
}

__device__ __forceinline__ Hfv hfv_at(const StepArgs& a, int dir, long long c) {
  return hfv_fwd(dir, a.U.h[c], a.U.hu[c], a.U.hv[c], a.kact[c], a.S.gz[c], a.tiny);
}

// Face records of one tile: every face flux and every one-sided difference
// the tile's cells use is computed once (a face is shared by two cells) and
// staged in LDS.  x-face fi of row fj lies between tile columns fi-1 and fi.
template <int TX>
struct FaceLds {
  static constexpr int TY = BLOCK / TX, RX = TX + 2, RXY = (TX + 2) * (TY + 2);
  double fx[3][TY][TX + 1], fy[3][TY + 1][TX];
  double dxh[TY][TX + 1], dxhv[TY][TX + 1];  // (U_right - U_left) / dx of h, hv
  double dyh[TY + 1][TX], dyhu[TY + 1][TX];  // (U_top - U_bottom) / dy of h, hu
  // records of the tile and its one-cell frame: the state and, for h >=
  // tiny, hu/h, hv/h and sqrt(kact h g2) -- hfv_fwd's values, evaluated
  // once per cell instead of once per face
  double h[RXY], hu[RXY], hv[RXY], vx[RXY], vy[RXY], sq[RXY];
};

// hfv_fwd (tg_math.h) of record r
template <int TX>
__device__ __forceinline__ Hfv rec_hfv(const FaceLds<TX>& L, int dir, int r, double tiny) {
  Hfv o;
  const double h = L.h[r], hu = L.hu[r], hv = L.hv[r];
  if (h < tiny) {
#pragma unroll
    for (int i = 0; i < 3; i++) o.u[i] = o.f[i] = o.w[i] = 0.;
    return o;
  }
  const double a = L.sq[r];
  o.u[0] = h; o.u[1] = hu; o.u[2] = hv;
  if (dir == 0) {
    const double vel = L.vx[r];
    o.f[0] = hu; o.f[1] = hu * vel + 0.5 * a * a * h; o.f[2] = hv * vel;
    o.w[0] = vel - a; o.w[1] = vel; o.w[2] = vel + a;
  } else {
    const double vel = L.vy[r];
    o.f[0] = hv; o.f[1] = hu * vel; o.f[2] = hv * vel + 0.5 * a * a * h;
    o.w[0] = vel - a; o.w[1] = vel; o.w[2] = vel + a;
  }
  return o;
}

// Update pass 1 (one workgroup per mark block, one cell of each of its four
// tiles per thread): a tile whose cells and side neighbours all hold
// h < tiny has every face flux zero (the physical flux is zero below tiny
// and riemannflux of two such sides is zero) and no source terms
// (update_states adds them only above tiny); its cells get update_states'
// own arithmetic with zero fluxes, bit for bit, here.  Every other tile is
// appended to the list pass 2 walks.  A tile the input marks show to hold
// +0.0 is not read (its cells are +0.0); when it is dry, its output tile
// already holds +0.0 and dt is finite, the formula gives +0.0 again and
// nothing is written.
template <int TX>
__global__ __launch_bounds__(BLOCK) void k_update_classify(StepArgs a) {
  constexpr int TY = BLOCK / TX;
  const BlockDev B = a.blocks[blockIdx.x];
  const double dt = a.dt_rec[(long long)B.inst * a.max_steps + a.step];
  if (a.mk_in && update_skip(a.mk_in, a.mk_out, B, dt)) return;
  const InstDev I = a.inst[B.inst];
  // tiles of the block holding +0.0 in the input / output state (mark bytes)
  const uint32_t zin = (a.mk_in && a.zbytes) ? a.mk_in[B.mw] : 0u;
  const uint32_t zout = (a.mk_out && a.zbytes && isfinite(dt)) ? a.mk_out[B.mw] : 0u;
  __shared__ unsigned s_wet, s_nz;
  if (threadIdx.x == 0) { s_wet = 0u; s_nz = 0u; }
  Cell q[4];
  block_cells<TX>(B, q);
  const long long nx = I.nx;
  const int tx = (int)(threadIdx.x % TX), ty = (int)(threadIdx.x / TX);
  unsigned wet = 0u;
  // h of every cell not known to be +0.0 (the wet test); the momenta only
  // once the tile is known to be dry
  double out[4][3];
#pragma unroll
  for (int b = 0; b < 4; b++) {
    if (!q[b].act) continue;
    const long long c = q[b].c;
    out[b][0] = out[b][1] = out[b][2] = 0.;
    if (!((zin >> (8 * b)) & 1u)) out[b][0] = a.U.h[c];
    bool w = !(out[b][0] < a.tiny);
    if (tx == 0 && q[b].hasW) w = w || !(a.U.h[c - 1] < a.tiny);
    if (tx == TX - 1 && q[b].hasE) w = w || !(a.U.h[c + 1] < a.tiny);
    if (ty == 0 && q[b].hasS) w = w || !(a.U.h[c - nx] < a.tiny);
    if (ty == TY - 1 && q[b].hasN) w = w || !(a.U.h[c + nx] < a.tiny);
    if (w) wet |= 1u << b;
  }
  __syncthreads();  // s_wet, s_nz set
  if (wet) atomicOr(&s_wet, wet);
  __syncthreads();
  const unsigned W = s_wet;
  if (threadIdx.x == 0 && W) {
    int k = atomicAdd(a.tile_count, __popc(W));
#pragma unroll
    for (int b = 0; b < 4; b++)
      if ((W >> b) & 1u) a.tile_list[k++] = B.tile[b];
  }
  // the momenta of the dry tiles that hold something (a wet tile's cells are
  // read again by the full path)
#pragma unroll
  for (int b = 0; b < 4; b++) {
    if (!q[b].act || ((W >> b) & 1u) || ((zin >> (8 * b)) & 1u)) continue;
    out[b][1] = a.U.hu[q[b].c];
    out[b][2] = a.U.hv[q[b].c];
  }
  const double dtdx = dt / I.dx, dtdy = dt / I.dy, zero = 0.;
  unsigned nz = 0u;
  int nshort = 0;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    if (!q[b].act || ((W >> b) & 1u)) continue;
    nshort++;
#pragma unroll
    for (int v = 0; v < 3; v++)
      out[b][v] = out[b][v] - dtdx * (zero - zero) - dtdy * (zero - zero);
    if (out[b][0] < 0.) out[b][0] = 0;
    if (!isfinite(out[b][0]) || !isfinite(out[b][1]) || !isfinite(out[b][2])) atomicOr(a.err, 1);
    if (!bitzero3(out[b][0], out[b][1], out[b][2])) nz |= 1u << b;
  }
  if (a.short_cells) {
    const int cnt = (int)wave_sum((double)nshort);
    if (threadIdx.x % 64 == 0 && cnt) atomicAdd(a.short_cells, (unsigned long long)cnt);
  }
  // the marks' reduction before the stores: no barrier waits on them
  if (a.mk_out) {
    if (nz) atomicOr(&s_nz, nz);
    __syncthreads();
  }
  // a dry tile of +0.0 whose output already holds +0.0 (finite dt): the
  // formula gives +0.0 again, nothing to write
  unsigned keep = 0u;
#pragma unroll
  for (int b = 0; b < 4; b++)
    if (((zin & zout) >> (8 * b)) & 1u) keep |= 1u << b;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    if (!q[b].act || (((W | keep) >> b) & 1u)) continue;
    a.h_out[q[b].c] = out[b][0];
    a.hu_out[q[b].c] = out[b][1];
    a.hv_out[q[b].c] = out[b][2];
  }
  if (a.mk_out && threadIdx.x == 0) {
    const unsigned NZ = s_nz;
#pragma unroll
    for (int b = 0; b < 4; b++)
      if (B.tile[b] >= 0 && !(((W | keep) >> b) & 1u))
        mark_byte_store(a.mk_out, B.mw, b, !((NZ >> b) & 1u));
  }
}

// update_tile's per-cell part: slopes from the staged differences, the four
// face fluxes with outflow copies, update_states; z: the new state is +0.0.
template <int TX>
__device__ __forceinline__ void update_cell(const StepArgs& a, const TileDev& t,
                                            const InstDev& I, const Cell& q, FaceLds<TX>& L,
                                            bool& z) {
  const int tx = (int)(threadIdx.x % TX), ty = (int)(threadIdx.x / TX);
  const long long c = q.c;
  const double dt = a.dt_rec[(long long)t.inst * a.max_steps + a.step];
  const double prev[3] = {a.U.h[c], a.U.hu[c], a.U.hv[c]};
  // Element::get_slopes, element2.C:830-939 (only the components the
  // corrector reads: hx, hvx, hy, huy); boundary cells keep zero slopes
  double hx = 0., hvx = 0., hy = 0., huy = 0.;
  if (q.hasE && q.hasW && q.hasN && q.hasS) {
    hx = slope_from(L.dxh[ty][tx + 1], L.dxh[ty][tx], I.dx, a.gamma, a.tiny);
    hvx = slope_from(L.dxhv[ty][tx + 1], L.dxhv[ty][tx], I.dx, a.gamma, a.tiny);
    hy = slope_from(L.dyh[ty + 1][tx], L.dyh[ty][tx], I.dy, a.gamma, a.tiny);
    huy = slope_from(L.dyhu[ty + 1][tx], L.dyhu[ty][tx], I.dy, a.gamma, a.tiny);
  }
  // a face on the domain boundary takes the flux of the opposite face
  // (boundary_flux, element2.C:3374-3411)
  double fE[3], fW[3], fN[3], fS[3];
  for (int v = 0; v < 3; v++) {
    fE[v] = q.hasE ? L.fx[v][ty][tx + 1] : (q.hasW ? L.fx[v][ty][tx] : 0.);
    fW[v] = q.hasW ? L.fx[v][ty][tx] : fE[v];
    fN[v] = q.hasN ? L.fy[v][ty + 1][tx] : (q.hasS ? L.fy[v][ty][tx] : 0.);
    fS[v] = q.hasS ? L.fy[v][ty][tx] : fN[v];
  }
  // correct() -> update_states (corrector.C:36-140, residual.C:162-275)
  const double kc = (prev[0] > a.tiny) ? a.kact[c] : I.eps;
  double out[3];
  SrcInfo si;
  update_states(prev, out, fE, fN, fW, fS, dt / I.dx, dt / I.dy, dt, hx, hvx, hy, huy,
                a.S.kx[c], a.S.ky[c], I.sin_int, mat_of(I, a.mat, a.mtab, c).tan_bed,
                a.S.gx[c], a.S.gy[c], a.S.gz[c],
                a.S.dgx[c], a.S.dgy[c], kc, a.ft, a.tiny, &si);
  if (!isfinite(out[0]) || !isfinite(out[1]) || !isfinite(out[2])) atomicOr(a.err, 1);
  a.h_out[c] = out[0];
  a.hu_out[c] = out[1];
  a.hv_out[c] = out[2];
  z = bitzero3(out[0], out[1], out[2]);
}

// Update pass 2, one listed tile: face fluxes staged in LDS, slopes,
// update_states.
template <int TX>
__device__ __forceinline__ void update_tile(const StepArgs& a, int tile, FaceLds<TX>& L) {
  constexpr int TY = BLOCK / TX;
  const TileDev t = a.tiles[tile];
  const InstDev I = a.inst[t.inst];
  const Cell q = cell_of<TX>(t, I);
  const long long nx = I.nx;
  if (a.active) {
    const int cnt = __syncthreads_count(q.act);
    if (threadIdx.x == 0) atomicAdd(a.active, (unsigned long long)cnt);
  }
  // phase 0: records of the tile and its frame (cells of the member only)
  constexpr int RX = FaceLds<TX>::RX, RXY = FaceLds<TX>::RXY;
  for (int r = (int)threadIdx.x; r < RXY; r += BLOCK) {
    const int gi = t.x0 - 1 + r % RX, gj = t.y0 - 1 + r / RX;
    if (gi >= 0 && gi < I.nx && gj >= 0 && gj < I.ny) {
      const long long c = I.off + (long long)gj * nx + gi;
      const double h = a.U.h[c], hu = a.U.hu[c], hv = a.U.hv[c];
      L.h[r] = h; L.hu[r] = hu; L.hv[r] = hv;
      double vx = 0., vy = 0., sq = 0.;
      if (!(h < a.tiny)) {
        vx = hu / h;
        vy = hv / h;
        sq = sqrt(a.kact[c] * h * a.S.gz[c]);
      }
      L.vx[r] = vx; L.vy[r] = vy; L.sq[r] = sq;
    }
  }
  __syncthreads();
  // phase 1: faces with a cell of the member on both sides
  // (calc_edge_states, element2.C:2744-3072)
  for (int f = (int)threadIdx.x; f < (TX + 1) * TY; f += BLOCK) {
    const int fi = f % (TX + 1), fj = f / (TX + 1);
    const int gi = t.x0 + fi, gj = t.y0 + fj;
    if (gi > 0 && gi < I.nx && gj < I.ny) {
      const int r = (fj + 1) * RX + fi + 1;  // record of the right cell
      double fl[3];
      riemann(rec_hfv<TX>(L, 0, r - 1, a.tiny), rec_hfv<TX>(L, 0, r, a.tiny), fl);
      for (int v = 0; v < 3; v++) L.fx[v][fj][fi] = fl[v];
      L.dxh[fj][fi] = (L.h[r] - L.h[r - 1]) / I.dx;
      L.dxhv[fj][fi] = (L.hv[r] - L.hv[r - 1]) / I.dx;
    }
  }
  for (int f = (int)threadIdx.x; f < TX * (TY + 1); f += BLOCK) {
    const int fi = f % TX, fj = f / TX;
    const int gi = t.x0 + fi, gj = t.y0 + fj;
    if (gj > 0 && gj < I.ny && gi < I.nx) {
      const int r = (fj + 1) * RX + fi + 1;  // record of the upper cell
      double fl[3];
      riemann(rec_hfv<TX>(L, 1, r - RX, a.tiny), rec_hfv<TX>(L, 1, r, a.tiny), fl);
      for (int v = 0; v < 3; v++) L.fy[v][fj][fi] = fl[v];
      L.dyh[fj][fi] = (L.h[r] - L.h[r - RX]) / I.dy;
      L.dyhu[fj][fi] = (L.hu[r] - L.hu[r - RX]) / I.dy;
    }
  }
  __syncthreads();
  bool z = true;
  if (q.act) update_cell<TX>(a, t, I, q, L, z);
  if (a.mk_out) mark_store<TX>(a.mk_out, t, __syncthreads_and(z));
}


template <int TX>
__global__ __launch_bounds__(BLOCK) void k_update(StepArgs a) {
  __shared__ FaceLds<TX> L;
  const int n = *a.tile_count;
  for (int i = (int)blockIdx.x; i < n; i += (int)gridDim.x) {
    update_tile<TX>(a, a.tile_list[list_band(i, n)], L);
    __syncthreads();  // LDS reused by the next tile
  }
}

__global__ __launch_bounds__(BLOCK) void k_stats(const InstDev* inst, const TileDev* tiles,
                                                 Fields U, const uint32_t* mk,
                                                 double* part_mass, double* part_hmax,
                                                 double tiny) {
  const TileDev t = tiles[blockIdx.x];
  if (mk && mk[t.mw] == MARK_FULL) {  // the block holds +0.0: no mass, no depth
    if (threadIdx.x == 0) { part_mass[blockIdx.x] = 0.; part_hmax[blockIdx.x] = 0.; }
    return;
  }
  const InstDev I = inst[t.inst];
  const Cell q = cell_of<64>(t, I);
  double m = 0., hm = 0.;
  if (q.act) {
    double h = U.h[q.c];
    m = I.dx * I.dy * h;
    if (h > tiny) hm = h;
  }
  __shared__ double sm[BLOCK / 64], sh[BLOCK / 64];
  m = wave_sum(m);
  hm = wave_max(hm);
  if (threadIdx.x % 64 == 0) { sm[threadIdx.x / 64] = m; sh[threadIdx.x / 64] = hm; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < BLOCK / 64; k++) { m += sm[k]; hm = (sh[k] > hm) ? sh[k] : hm; }
    part_mass[blockIdx.x] = m;
    part_hmax[blockIdx.x] = hm;
  }
}

// Copies a state buffer's three fields (raw bits, W the element's width)
// one mark block per workgroup and gives the source and the copy the zero
// marks of what it read: byte b of the block's word is 1 when tile b
// (block_cells' order, TX the update's tile width) holds +0.0 in all three
// fields.  A source whose marks were cleared (a state set from the host)
// and its checkpoint copy then let the next steps skip what is empty.  With
// dst and mk_dst null, only the source's marks are written.  A block of the
// copy marked whole +0.0 (MARK_FULL, meaningful in every tile shape) whose
// source block is +0.0 is not rewritten.
template <int TX, class W>
__global__ __launch_bounds__(BLOCK) void k_copy_marking(const BlockDev* blocks, const W* s0,
                                                        const W* s1, const W* s2, W* dst,
                                                        long long ncell, uint32_t* mk_src,
                                                        uint32_t* mk_dst) {
  const BlockDev B = blocks[blockIdx.x];
  __shared__ unsigned s_nz;
  if (threadIdx.x == 0) s_nz = 0u;
  // a copy whose block already holds +0.0 (word MARK_FULL) is not rewritten
  // where the source block is +0.0 too
  const bool dst_zero = mk_dst && mk_dst[B.mw] == MARK_FULL;
  Cell q[4];
  block_cells<TX>(B, q);
  W v[4][3];
  unsigned nz = 0u;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    v[b][0] = v[b][1] = v[b][2] = (W)0;
    if (!q[b].act) continue;
    const long long c = q[b].c;
    v[b][0] = s0[c]; v[b][1] = s1[c]; v[b][2] = s2[c];
    if ((v[b][0] | v[b][1] | v[b][2]) != (W)0) nz |= 1u << b;
  }
  __syncthreads();  // s_nz cleared
  if (nz) atomicOr(&s_nz, nz);
  __syncthreads();
  if (dst && !(dst_zero && s_nz == 0u)) {
#pragma unroll
    for (int b = 0; b < 4; b++) {
      if (!q[b].act) continue;
      const long long c = q[b].c;
      dst[c] = v[b][0];
      dst[ncell + c] = v[b][1];
      dst[2 * ncell + c] = v[b][2];
    }
  }
  if (threadIdx.x == 0) {
    uint32_t w = 0u;
#pragma unroll
    for (int b = 0; b < 4; b++)
      if (B.tile[b] >= 0 && !((s_nz >> b) & 1u)) w |= 1u << (8 * b);
    mk_src[B.mw] = w;
    if (mk_dst) mk_dst[B.mw] = w;
  }
}

// One workgroup per member b: out[b] = sum of sum_part over member b's
// tiles idx[off[b] .. off[b+1]) (each lane adds its strided share in list
// order, the lanes combine by a fixed tree, so the result does not depend on
// the launch), out[n + b] = the max of max_part over them (when given),
// out[2n + b] = time[b] and out[3n + b] = dt_rec[b * stride + col] (when
// given).  Replaces a copy of every tile's partial and a host loop over them.
__global__ __launch_bounds__(BLOCK) void k_member_reduce(const double* sum_part,
                                                         const double* max_part,
                                                         const int* off, const int* idx, int n,
                                                         const double* time,
                                                         const double* dt_rec, long long stride,
                                                         int col, double* out) {
  const int b = blockIdx.x;
  double s = 0., m = 0.;
  for (int j = off[b] + (int)threadIdx.x; j < off[b + 1]; j += BLOCK) {
    const int t = idx[j];
    s += sum_part[t];
    if (max_part) m = (max_part[t] > m) ? max_part[t] : m;
  }
  __shared__ double ss[BLOCK / 64], sm[BLOCK / 64];
  s = wave_sum(s);
  m = wave_max(m);
  if (threadIdx.x % 64 == 0) { ss[threadIdx.x / 64] = s; sm[threadIdx.x / 64] = m; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < BLOCK / 64; k++) { s += ss[k]; m = (sm[k] > m) ? sm[k] : m; }
    out[b] = s;
    if (max_part) out[n + b] = m;
    if (time) out[2 * n + b] = time[b];
    if (dt_rec) out[3 * n + b] = dt_rec[(long long)b * stride + col];
  }
}

// ---------------------------------------------------------------- adjoint --

// Everything the Jacobian blocks need about one cell at U^s
// (DualElem after update_dual_grid + update_states in calc_jacobian).
struct DualInfo {
  double h, hu, hv, g2;
  double k;              // kact from the dual slopes (dual_grid_routines.C:68-84)
  double dk;             // d kact / d phi_bed
  double d[6];           // dual slopes hx, hux, hvx, hy, huy, hvy
  double hsx[3], hsy[3]; // hslope_sens(dir, self/minus/plus)
  SrcInfo si;
};

__device__ __forceinline__ DualInfo dual_info(const AdjArgs& a, const InstDev& I, int i, int j,
                                              bool want_dk) {
  DualInfo r;
  const long long nx = I.nx, c = I.off + (long long)j * nx + i;
  r.h = a.U.h[c]; r.hu = a.U.hu[c]; r.hv = a.U.hv[c]; r.g2 = a.S.gz[c];
#pragma unroll
  for (int k = 0; k < 6; k++) r.d[k] = 0.;
#pragma unroll
  for (int k = 0; k < 3; k++) r.hsx[k] = r.hsy[k] = 0.;
  if (i + 1 < I.nx && i > 0 && j + 1 < I.ny && j > 0) {
    const double ix = I.inv_dx, isx = I.inv_sx;
    const double iy = I.inv_dy, isy = I.inv_sy;
    DualSlope s = slope_dual(a.U.h[c + 1], r.h, a.U.h[c - 1], I.dx, ix, isx, a.gamma, a.tiny, true);
    r.d[0] = s.d; r.hsx[0] = s.s0; r.hsx[1] = s.s1; r.hsx[2] = s.s2;
    r.d[1] = slope_dual(a.U.hu[c + 1], r.hu, a.U.hu[c - 1], I.dx, ix, isx, a.gamma, a.tiny, false).d;
    r.d[2] = slope_dual(a.U.hv[c + 1], r.hv, a.U.hv[c - 1], I.dx, ix, isx, a.gamma, a.tiny, false).d;
    s = slope_dual(a.U.h[c + nx], r.h, a.U.h[c - nx], I.dy, iy, isy, a.gamma, a.tiny, true);
    r.d[3] = s.d; r.hsy[0] = s.s0; r.hsy[1] = s.s1; r.hsy[2] = s.s2;
    r.d[4] = slope_dual(a.U.hu[c + nx], r.hu, a.U.hu[c - nx], I.dy, iy, isy, a.gamma, a.tiny, false).d;
    r.d[5] = slope_dual(a.U.hv[c + nx], r.hv, a.U.hv[c - nx], I.dy, iy, isy, a.gamma, a.tiny, false).d;
  }
  const MatDev M = mat_of(I, a.mat, a.mtab, c);
  r.k = getcoef(r.h, r.hu, r.hv, r.d[0], r.d[1], r.d[3], r.d[5], I.kc1, M.kroot, I.eps, a.tiny);
  r.dk = want_dk ? dgetcoef_dphi(r.h, r.hu, r.hv, r.d[0], r.d[1], r.d[3], r.d[5], I.kc1, M.kroot,
                                 M.dkroot, I.eps, a.tiny)
                 : 0.;
  r.si = src_switches(r.h, r.hu, r.hv, r.d[0], r.d[2], r.d[3], r.d[4], a.S.kx[c], a.S.ky[c],
                      M.tan_bed, r.g2, a.ft, a.tiny);
  return r;
}

// J = dtdx*(px - nx) + dtdy*(py - ny), calc_jacobian_elem
// (compute_jacobian_elem.C:21-24); an operand whose flag is false is
// ZERO_MATRIX.  Operands are taken by value and selected element by element
// so the blocks stay in registers.
__device__ __forceinline__ M3 flux_block(const M3& px, bool ux, const M3& nxm, bool unx,
                                         const M3& py, bool uy, const M3& nym, bool uny,
                                         double dtdx, double dtdy) {
  M3 J;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double a = ux ? px.a[i][j] : 0., b = unx ? nxm.a[i][j] : 0.;
      double c = uy ? py.a[i][j] : 0., d = uny ? nym.a[i][j] : 0.;
      J.a[i][j] = +dtdx * (a - b) + dtdy * (c - d);
    }
  return J;
}

// Source-term part of a neighbour block (compute_jacobian_elem.C:44-51).
__device__ __forceinline__ void nbr_source(M3& J, const DualInfo& n, double sin_int, double dt,
                                           double dh0, double dh1, double tiny) {
  if (n.h > tiny) {
    double alpha = sin_int * n.si.sgn[0] * n.k;
    double betta = sin_int * n.si.sgn[1] * n.k;
    if (!n.si.stop[0]) J.a[1][0] -= -dt * n.h * alpha * n.g2 * dh1;
    if (!n.si.stop[1]) J.a[2][0] -= -dt * n.h * betta * n.g2 * dh0;
  }
}

__device__ __forceinline__ void accum_T(double adj[3], const M3& J, const double lam[3]) {
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int l = 0; l < 3; l++) adj[k] += lam[l] * J.a[l][k];
}


// The part of a cell's dual record its neighbours read, staged in LDS for the
// tile plus a one-cell halo: each record is computed once per workgroup
// instead of once per reading cell.
// LDS record slots: the largest tile plus halo over the built tile widths
// ((TX + 2) * (BLOCK / TX + 2) is largest at TX = 64)
constexpr int RN = (64 + 2) * (BLOCK / 64 + 2);
struct DualLds {
  double h[RN], hu[RN], hv[RN], g2[RN], k[RN], dk[RN];
  double hsx1[RN], hsx2[RN], hsy1[RN], hsy2[RN];
  int flags[RN];  // sgn0+1 | (sgn1+1)<<2 | stop0<<4 | stop1<<5
  // per record, for h >= tiny: 1/h and sqrt(k h g2) (dual_hfv, hfv_fwd,
  // dhfv_dk), hu/h and hv/h (hfv_fwd), dhfv_dk's da -- each evaluated once
  // with the expression the flux functions use, instead of once per face
  double inv[RN], sq[RN], vx[RN], vy[RN], da[RN];
};

__device__ __forceinline__ void lds_put(DualLds& L, int r, const DualInfo& d, double tiny) {
  L.h[r] = d.h; L.hu[r] = d.hu; L.hv[r] = d.hv; L.g2[r] = d.g2; L.k[r] = d.k; L.dk[r] = d.dk;
  double inv = 0., sq = 0., vx = 0., vy = 0., da = 0.;
  if (!(d.h < tiny)) {
    inv = 1. / d.h;
    sq = sqrt(d.k * d.h * d.g2);
    vx = d.hu / d.h;
    vy = d.hv / d.h;
    da = (sq > 0.) ? (d.h * d.g2) / (2. * sq) : 0.;
  }
  L.inv[r] = inv; L.sq[r] = sq; L.vx[r] = vx; L.vy[r] = vy; L.da[r] = da;
  L.hsx1[r] = d.hsx[1]; L.hsx2[r] = d.hsx[2]; L.hsy1[r] = d.hsy[1]; L.hsy2[r] = d.hsy[2];
  L.flags[r] = (int)(d.si.sgn[0] + 1.) | ((int)(d.si.sgn[1] + 1.) << 2) | (d.si.stop[0] << 4) |
               (d.si.stop[1] << 5);
}

// Rebuilds the fields a neighbour block and the face Jacobians use; the
// other fields (own slopes, sgn[2..3]) belong to the owning cell only.
__device__ __forceinline__ DualInfo lds_get(const DualLds& L, int r) {
  DualInfo d;
  d.h = L.h[r]; d.hu = L.hu[r]; d.hv = L.hv[r]; d.g2 = L.g2[r]; d.k = L.k[r]; d.dk = L.dk[r];
  d.hsx[0] = d.hsy[0] = 0.;
  d.hsx[1] = L.hsx1[r]; d.hsx[2] = L.hsx2[r]; d.hsy[1] = L.hsy1[r]; d.hsy[2] = L.hsy2[r];
  const int f = L.flags[r];
  d.si.sgn[0] = (double)((f & 3) - 1);
  d.si.sgn[1] = (double)(((f >> 2) & 3) - 1);
  d.si.sgn[2] = d.si.sgn[3] = 0.;
  d.si.stop[0] = (f >> 4) & 1;
  d.si.stop[1] = (f >> 5) & 1;
#pragma unroll
  for (int k = 0; k < 6; k++) d.d[k] = 0.;
  return d;
}

// dual_hfv of record r (tg_math.h), with its staged 1/h and sqrt(k h g2)
__device__ __forceinline__ DualHfv lds_dual_hfv(const DualLds& L, int dir, int r, double tiny) {
  DualHfv o;
  const double h = L.h[r], hu = L.hu[r], hv = L.hv[r];
  if (h < tiny) {
#pragma unroll
    for (int i = 0; i < 3; i++) o.u[i] = o.f[i] = o.w[i] = 0.;
    o.fj = zero3(); o.sj = zero3();
    return o;
  }
  const double h_inv = L.inv[r], a = L.sq[r];
  o.u[0] = h; o.u[1] = hu; o.u[2] = hv;
  if (dir == 0) {
    double vel = hu * h_inv;
    o.f[0] = hu; o.f[1] = hu * vel + 0.5 * a * a * h; o.f[2] = hv * vel;
    o.w[0] = vel - a; o.w[1] = vel; o.w[2] = vel + a;
    o.fj.a[0][0] = 0.; o.fj.a[0][1] = 1.; o.fj.a[0][2] = 0.;
    o.fj.a[1][0] = a * a - vel * vel; o.fj.a[1][1] = 2 * vel; o.fj.a[1][2] = 0.;
    o.fj.a[2][0] = -vel * hv * h_inv; o.fj.a[2][1] = hv * h_inv; o.fj.a[2][2] = vel;
    o.sj.a[0][0] = -.5 * a * h_inv - vel * h_inv; o.sj.a[0][1] = h_inv; o.sj.a[0][2] = 0.;
    o.sj.a[1][0] = -vel * h_inv; o.sj.a[1][1] = h_inv; o.sj.a[1][2] = 0.;
    o.sj.a[2][0] = .5 * a * h_inv - vel * h_inv; o.sj.a[2][1] = h_inv; o.sj.a[2][2] = 0.;
  } else {
    double vel = hv * h_inv;
    o.f[0] = hv; o.f[1] = hu * vel; o.f[2] = hv * vel + 0.5 * a * a * h;
    o.w[0] = vel - a; o.w[1] = vel; o.w[2] = vel + a;
    o.fj.a[0][0] = 0.; o.fj.a[0][1] = 0.; o.fj.a[0][2] = 1.;
    o.fj.a[1][0] = -vel * hu * h_inv; o.fj.a[1][1] = vel; o.fj.a[1][2] = hu * h_inv;
    o.fj.a[2][0] = a * a - vel * vel; o.fj.a[2][1] = 0.; o.fj.a[2][2] = 2 * vel;
    o.sj.a[0][0] = -.5 * a * h_inv - vel * h_inv; o.sj.a[0][1] = 0.; o.sj.a[0][2] = h_inv;
    o.sj.a[1][0] = -vel * h_inv; o.sj.a[1][1] = 0.; o.sj.a[1][2] = h_inv;
    o.sj.a[2][0] = .5 * a * h_inv - vel * h_inv; o.sj.a[2][1] = 0.; o.sj.a[2][2] = h_inv;
  }
  return o;
}

// hfv_fwd and dhfv_dk (tg_math.h) of record r, with its staged values
__device__ __forceinline__ Hfv lds_hfv(const DualLds& L, int dir, int r, double tiny) {
  Hfv o;
  const double h = L.h[r], hu = L.hu[r], hv = L.hv[r];
  if (h < tiny) {
#pragma unroll
    for (int i = 0; i < 3; i++) o.u[i] = o.f[i] = o.w[i] = 0.;
    return o;
  }
  const double a = L.sq[r];
  o.u[0] = h; o.u[1] = hu; o.u[2] = hv;
  if (dir == 0) {
    const double vel = L.vx[r];
    o.f[0] = hu; o.f[1] = hu * vel + 0.5 * a * a * h; o.f[2] = hv * vel;
    o.w[0] = vel - a; o.w[1] = vel; o.w[2] = vel + a;
  } else {
    const double vel = L.vy[r];
    o.f[0] = hv; o.f[1] = hu * vel; o.f[2] = hv * vel + 0.5 * a * a * h;
    o.w[0] = vel - a; o.w[1] = vel; o.w[2] = vel + a;
  }
  return o;
}
__device__ __forceinline__ void lds_dhfv_dk(const DualLds& L, int dir, int r, double tiny,
                                            double dF[3], double dW[3]) {
  dF[0] = dF[1] = dF[2] = dW[0] = dW[1] = dW[2] = 0.;
  const double h = L.h[r];
  if (h < tiny) return;
  const double g3 = L.g2[r], da = L.da[r];
  if (dir == 0) dF[1] = 0.5 * h * g3 * h;
  else dF[2] = 0.5 * h * g3 * h;
  dW[0] = -da;
  dW[2] = da;
}

// face flux and its derivative along d kact/d phi_bed (both sides) between
// records rl and rr
__device__ __forceinline__ void face_dk(const AdjArgs& a, int side, const DualLds& L, int rl,
                                        int rr, double flux[3], double dflux[3]) {
  Hfv hl = lds_hfv(L, side, rl, a.tiny);
  Hfv hr = lds_hfv(L, side, rr, a.tiny);
  riemann(hl, hr, flux);
  double dFl[3], dWl[3], dFr[3], dWr[3];
  lds_dhfv_dk(L, side, rl, a.tiny, dFl, dWl);
  lds_dhfv_dk(L, side, rr, a.tiny, dFr, dWr);
  const double dkl = L.dk[rl], dkr = L.dk[rr];
#pragma unroll
  for (int v = 0; v < 3; v++) {
    dFl[v] *= dkl; dWl[v] *= dkl; dFr[v] *= dkr; dWr[v] *= dkr;
  }
  driemann(hl, hr, dFl, dWl, dFr, dWr, dflux);
}

// lambda^{s+1}_c . d U^{s+1}_c / d phi_bed at frozen dt (bed-friction
// gradient; no reference counterpart).  C: the cell's own record; rc, rE,
// rN, rW, rS: the LDS records of the cell and its neighbours (absent
// neighbours' slots are passed but not read).  Faces with no neighbour copy
// the opposite face (outflow), as in the forward step.
__device__ double dphi_term(const AdjArgs& a, const InstDev& I, const Cell& q,
                            const DualInfo& C, const DualLds& L, int rc, int rE, int rN, int rW,
                            int rS, double dt, const double lam[3]) {
  double fE[3] = {0., 0., 0.}, fN[3] = {0., 0., 0.}, fW[3] = {0., 0., 0.}, fS[3] = {0., 0., 0.};
  double dE[3] = {0., 0., 0.}, dN[3] = {0., 0., 0.}, dW[3] = {0., 0., 0.}, dS[3] = {0., 0., 0.};
  if (q.hasE) face_dk(a, 0, L, rc, rE, fE, dE);
  if (q.hasW) face_dk(a, 0, L, rW, rc, fW, dW);
  if (q.hasN) face_dk(a, 1, L, rc, rN, fN, dN);
  if (q.hasS) face_dk(a, 1, L, rS, rc, fS, dS);
#pragma unroll
  for (int v = 0; v < 3; v++) {
    if (!q.hasE) { fE[v] = fW[v]; dE[v] = dW[v]; }
    if (!q.hasW) { fW[v] = fE[v]; dW[v] = dE[v]; }
    if (!q.hasN) { fN[v] = fS[v]; dN[v] = dS[v]; }
    if (!q.hasS) { fS[v] = fN[v]; dS[v] = dN[v]; }
  }
  const double dtdx = dt / I.dx, dtdy = dt / I.dy;
  double out[3];
#pragma unroll
  for (int v = 0; v < 3; v++) out[v] = -dtdx * (dE[v] - dW[v]) - dtdy * (dN[v] - dS[v]);
  if (C.h - dtdx * (fE[0] - fW[0]) - dtdy * (fN[0] - fS[0]) < 0.) out[0] = 0.;
  if (C.h > a.tiny) {
    const long long c = q.c;
    const double tbc = mat_of(I, a.mat, a.mtab, c).tan_bed;
    const double sec2 = 1. + tbc * tbc;
    const double vx = L.vx[rc], vy = L.vy[rc];  // hu / h, hv / h (staged)
    const double g2 = C.g2, dg0 = a.S.dgx[c], dg1 = a.S.dgy[c], k0 = a.S.kx[c], k1 = a.S.ky[c];
    double ds2x = C.si.sgn[0] * C.h * C.dk * (g2 * C.d[3] + dg1 * C.h) * I.sin_int;
    double ds3x = C.si.sgn[2] * umax(g2 * C.h + vx * C.hu * k0, 0.0) * sec2;
    double ds2y = C.si.sgn[1] * C.h * C.dk * (g2 * C.d[0] + dg0 * C.h) * I.sin_int;
    double ds3y = C.si.sgn[3] * umax(g2 * C.h + vy * C.hv * k1, 0.0) * sec2;
    out[1] += dt * (-ds2x - ds3x);
    out[2] += dt * (-ds2y - ds3y);
  }
  return lam[0] * out[0] + lam[1] * out[1] + lam[2] * out[2];
}

// Adjoint pass 1 (one workgroup per mark block, one cell of each of its four
// tiles per thread): a tile of void cells (h == 0) gets lambda^s = 0 --
// every block that multiplies a void cell's state is zero
// (calc_jacobain.C:117-122), as in the per-cell path of pass 2 -- and, with
// lambda^{s+1} == 0 on the tile, zero bed-friction terms; it is written
// here.  Every other tile is appended to the list pass 2 walks.
template <int TX>
__global__ __launch_bounds__(BLOCK) void k_adjoint_classify(AdjArgs a) {
  // the block record is read alongside the mark words (a skipped block
  // wastes its 64 bytes; a listed one saves a round trip)
  const BlockDev B = a.blocks[blockIdx.x];
  if (a.mk_out &&
      adjoint_skip(a.mk_U, a.dphi_part ? a.mk_lnext : nullptr, a.mk_out, (int)blockIdx.x))
    return;
  __shared__ unsigned s_busy;
  if (threadIdx.x == 0) s_busy = 0u;
  // tiles the marks show to be void (U +0.0 and, with the gradient,
  // lambda^{s+1} +0.0): nothing to read; output tiles already +0.0: nothing
  // to write for a void tile
  uint32_t zvoid = 0u, zout = 0u;
  if (a.mk_out && a.l_bytes) {
    if (a.u_bytes) zvoid = a.mk_U[B.mw];
    if (a.dphi_part) zvoid &= a.mk_lnext[B.mw];
    zout = a.mk_out[B.mw];
  }
  Cell q[4];
  block_cells<TX>(B, q);
  unsigned busy = 0u;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    if (!q[b].act || ((zvoid >> (8 * b)) & 1u)) continue;
    const long long c = q[b].c;
    bool u = a.U.h[c] != 0.;
    if (a.dphi_part)
      u = u || !(a.lam_next.h[c] == 0. && a.lam_next.hu[c] == 0. && a.lam_next.hv[c] == 0.);
    if (u) busy |= 1u << b;
  }
  __syncthreads();  // s_busy set
  if (busy) atomicOr(&s_busy, busy);
  __syncthreads();
  const unsigned W = s_busy;
  if (threadIdx.x == 0 && W) {
    int k = atomicAdd(a.tile_count, __popc(W));
#pragma unroll
    for (int b = 0; b < 4; b++)
      if ((W >> b) & 1u) a.tile_list[k++] = B.tile[b];
  }
  int nshort = 0;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    if (!q[b].act || ((W >> b) & 1u)) continue;
    nshort++;
    if ((zout >> (8 * b)) & 1u) continue;  // already +0.0
    a.lam_h[q[b].c] = 0.;
    a.lam_hu[q[b].c] = 0.;
    a.lam_hv[q[b].c] = 0.;
  }
  if (a.short_cells) {
    const int cnt = (int)wave_sum((double)nshort);
    if (threadIdx.x % 64 == 0 && cnt) atomicAdd(a.short_cells, (unsigned long long)cnt);
  }
  if (a.mk_out && threadIdx.x == 0) {
#pragma unroll
    for (int b = 0; b < 4; b++)
      if (B.tile[b] >= 0 && !((W >> b) & 1u)) mark_byte_store(a.mk_out, B.mw, b, true);
  }
}


// Adjoint pass 2, one listed tile.
template <int TX>
__device__ __forceinline__ void adjoint_tile(const AdjArgs& a, int tile, DualLds& L,
                                             double* sp) {
  const TileDev t = a.tiles[tile];
  const InstDev I = a.inst[t.inst];
  const Cell q = cell_of<TX>(t, I);
  double dphi = 0.;
  bool zero_out = true;
  const bool want_dphi = a.dphi_part != nullptr;
  // the halo is the four side strips: the 5-point stencil reads no corner
  constexpr int TY = BLOCK / TX, RX = TX + 2, RY = TY + 2, NHALO = 2 * TX + 2 * TY;
  static_assert(RX * RY <= RN && NHALO <= BLOCK, "tile does not fit the LDS records");
  if (a.active) {
    const int cnt = __syncthreads_count(q.act);
    if (threadIdx.x == 0) atomicAdd(a.active, (unsigned long long)cnt);
  }
  const int tx = (int)(threadIdx.x % TX), ty = (int)(threadIdx.x / TX);
  const int rself = (ty + 1) * RX + tx + 1;
  // phase 1: dual records of the tile (kept in registers) and of the halo
  DualInfo self;
  if (q.act) {
    self = dual_info(a, I, q.i, q.j, want_dphi);
    lds_put(L, rself, self, a.tiny);
  }
  if ((int)threadIdx.x < NHALO) {
    const int h = (int)threadIdx.x;
    int rx, ry;
    if (h < TX) { rx = h + 1; ry = 0; }
    else if (h < 2 * TX) { rx = h - TX + 1; ry = RY - 1; }
    else if (h < 2 * TX + TY) { rx = 0; ry = h - 2 * TX + 1; }
    else { rx = RX - 1; ry = h - 2 * TX - TY + 1; }
    const int gi = t.x0 - 1 + rx, gj = t.y0 - 1 + ry;
    if (gi >= 0 && gi < I.nx && gj >= 0 && gj < I.ny)
      lds_put(L, ry * RX + rx, dual_info(a, I, gi, gj, want_dphi), a.tiny);
  }
  __syncthreads();
  if (q.act) {
    const long long c = q.c, nx = I.nx;
    const double dt = a.dt_rec[(long long)t.inst * a.max_steps + a.step];
    const double dtdx = dt / I.dx, dtdy = dt / I.dy;
    const double hc = self.h;
    double adj[3] = {0., 0., 0.};
    // slots: 0 self, 1 E, 2 N, 3 W, 4 S
    const bool has[5] = {true, q.hasE, q.hasN, q.hasW, q.hasS};
    const long long nb[5] = {c, c + 1, c + nx, c - 1, c - nx};
    const int rn[5] = {rself, rself + 1, rself + RX, rself - 1, rself - RX};
    // Neighbour records are re-read from LDS at each use (not held in
    // registers across the phases); the empty asm statements keep the
    // compiler from merging the re-reads into long-lived registers.
    auto nbr_hfv = [&](int dir, int r) { return lds_dual_hfv(L, dir, r, a.tiny); };
    if (hc != 0.) {  // a void cell receives nothing (calc_jacobain.C:117-122)
      // c-side Jacobians of the four faces (dual_riemannflux)
      M3 JE = zero3(), JN = zero3(), JW = zero3(), JS = zero3();
      {
        const DualHfv cx = lds_dual_hfv(L, 0, rself, a.tiny);
        if (has[1]) JE = dual_riemann_side(cx, nbr_hfv(0, rn[1]), true);
        if (has[3]) JW = dual_riemann_side(nbr_hfv(0, rn[3]), cx, false);
      }
      {
        const DualHfv cy = lds_dual_hfv(L, 1, rself, a.tiny);
        if (has[2]) JN = dual_riemann_side(cy, nbr_hfv(1, rn[2]), true);
        if (has[4]) JS = dual_riemann_side(nbr_hfv(1, rn[4]), cy, false);
      }
      asm volatile("" ::: "memory");

      // ---- self block (effelement 0): fjc(.,.,0) with boundary copies
      {
        M3 J;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int j = 0; j < 3; j++) {
            const double pE = has[1] ? JE.a[i][j] : (has[3] ? JW.a[i][j] : 0.);
            const double nW = has[3] ? JW.a[i][j] : (has[1] ? JE.a[i][j] : 0.);
            const double pN = has[2] ? JN.a[i][j] : (has[4] ? JS.a[i][j] : 0.);
            const double nS = has[4] ? JS.a[i][j] : (has[2] ? JN.a[i][j] : 0.);
            J.a[i][j] = +dtdx * (pE - nW) + dtdy * (pN - nS);
          }
        for (int i = 0; i < 3; i++) J.a[i][i] -= 1.;
        const DualInfo& C = self;
        if (C.h > a.tiny) {
          const double alpha = I.sin_int * C.si.sgn[0] * C.k;
          const double betta = I.sin_int * C.si.sgn[1] * C.k;
          const double vx = L.vx[rself], vy = L.vy[rself];  // hu / h, hv / h (staged)
          const double vx_sq = vx * vx, vy_sq = vy * vy;
          const double tb = mat_of(I, a.mat, a.mtab, c).tan_bed;
          const double g0 = a.S.gx[c], g1 = a.S.gy[c], g2 = C.g2;
          const double dg0 = a.S.dgx[c], dg1 = a.S.dgy[c], k0 = a.S.kx[c], k1 = a.S.ky[c];
          const double dh0 = C.hsx[0], dh1 = C.hsy[0];
          J.a[1][0] -= dt * (g0 - alpha * C.h * (2 * dg1 + g2 * dh1) - alpha * g2 * C.d[3]);
          if (C.si.sgn[2] != 0. && !C.si.stop[0]) {
            J.a[1][0] -= -dt * C.si.sgn[2] * tb * (g2 - vx_sq * k0);
            J.a[1][1] -= -dt * C.si.sgn[2] * tb * (2 * vx * k0);
            J.a[1][2] -= 0.;
          }
          J.a[2][0] -= dt * (g1 - betta * C.h * (2 * dg0 + g2 * dh0) - betta * g2 * C.d[0]);
          if (C.si.sgn[3] != 0. && !C.si.stop[1]) {
            J.a[2][0] -= -dt * C.si.sgn[3] * tb * (g2 - vy_sq * k1);
            J.a[2][1] -= 0.;
            J.a[2][2] -= -dt * C.si.sgn[3] * tb * (2 * vy * k1);
          }
        }
        const double lam[3] = {a.lam_next.h[c], a.lam_next.hu[c], a.lam_next.hv[c]};
        accum_T(adj, J, lam);
      }
      // ---- neighbour blocks in the reference's visiting order S, E, N, W
      // (neighbor[0..3] with positive_x_side = 1).
      const int order[4] = {4, 1, 2, 3};
#pragma unroll
      for (int o = 0; o < 4; o++) {
        const int s = order[o];
        if (!has[s]) continue;
        const DualInfo N = lds_get(L, rn[s]);
        const long long n = nb[s];
        M3 J;
        if (s == 1) {         // c is the xm neighbour of n
          const bool bnd = !(q.i + 2 < I.nx);
          J = flux_block(JE, bnd, JE, true, JE, false, JE, false, dtdx, dtdy);
          nbr_source(J, N, I.sin_int, dt, N.hsx[1], 0., a.tiny);
        } else if (s == 3) {  // c is the xp neighbour of n
          const bool bnd = !(q.i - 2 >= 0);
          J = flux_block(JW, true, JW, bnd, JW, false, JW, false, dtdx, dtdy);
          nbr_source(J, N, I.sin_int, dt, N.hsx[2], 0., a.tiny);
        } else if (s == 2) {  // c is the ym neighbour of n
          const bool bnd = !(q.j + 2 < I.ny);
          J = flux_block(JN, false, JN, false, JN, bnd, JN, true, dtdx, dtdy);
          nbr_source(J, N, I.sin_int, dt, 0., N.hsy[1], a.tiny);
        } else {              // c is the yp neighbour of n
          const bool bnd = !(q.j - 2 >= 0);
          J = flux_block(JS, false, JS, false, JS, true, JS, bnd, dtdx, dtdy);
          nbr_source(J, N, I.sin_int, dt, 0., N.hsy[2], a.tiny);
        }
        const double lam[3] = {a.lam_next.h[n], a.lam_next.hu[n], a.lam_next.hv[n]};
        accum_T(adj, J, lam);
      }
    }
    double out[3];
    for (int j = 0; j < 3; j++) {
      double v = 0. - adj[j];
      if (fabs(v) < 1e-16) v = 0.;  // calc_adjoint.C:150-152
      out[j] = v;
    }
    a.lam_h[c] = out[0];
    a.lam_hu[c] = out[1];
    a.lam_hv[c] = out[2];
    zero_out = bitzero3(out[0], out[1], out[2]);
    if (want_dphi) {
      const double lam[3] = {a.lam_next.h[c], a.lam_next.hu[c], a.lam_next.hv[c]};
      asm volatile("" ::: "memory");
      // every rn[] slot lies inside the LDS tile; records of absent
      // neighbours are passed but not read (dphi_term checks q.has*).  A
      // cell with lambda^{s+1} = 0 adds lambda . dU/dphi = 0 (dU/dphi is
      // finite): not evaluated (the tile's sum may then hold +0.0 where the
      // full evaluation would give -0.0)
      if (!(lam[0] == 0. && lam[1] == 0. && lam[2] == 0.))
        dphi = dphi_term(a, I, q, self, L, rself, rn[1], rn[2], rn[3], rn[4], dt, lam);
    }
  }
  if (want_dphi) {
    dphi = wave_sum(dphi);
    if (threadIdx.x % 64 == 0) sp[threadIdx.x / 64] = dphi;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < BLOCK / 64; k++) dphi += sp[k];
      a.dphi_part[tile] += dphi;
    }
  }
  if (a.mk_out) mark_store<TX>(a.mk_out, t, __syncthreads_and(zero_out));
}


template <int TX>
__global__ __launch_bounds__(BLOCK) void k_adjoint(AdjArgs a) {
  __shared__ DualLds L;
  __shared__ double sp[BLOCK / 64];
  const int n = *a.tile_count;
  for (int i = (int)blockIdx.x; i < n; i += (int)gridDim.x) {
    adjoint_tile<TX>(a, a.tile_list[list_band(i, n)], L, sp);
    __syncthreads();  // LDS reused by the next tile
  }
}

}  // namespace

// one instantiation per built tile width; a.shape selects it
#define TG_LAUNCH_TX(shape, KERNEL, GRID, ...)                                      \
  do {                                                                             \
    switch (SHAPE_TX[shape]) {                                                     \
      case 64: hipLaunchKernelGGL(KERNEL<64>, GRID, dim3(BLOCK), 0, __VA_ARGS__); break; \
      case 32: hipLaunchKernelGGL(KERNEL<32>, GRID, dim3(BLOCK), 0, __VA_ARGS__); break; \
      default: hipLaunchKernelGGL(KERNEL<16>, GRID, dim3(BLOCK), 0, __VA_ARGS__); break; \
    }                                                                              \
  } while (0)

void launch_coef(const StepArgs& a, bool reduce, hipStream_t s) {
  TG_LAUNCH_TX(a.shape, k_coef, dim3(a.nblocks), s, a, reduce ? 1 : 0);
}
void launch_dt(const StepArgs& a, int ninst, hipStream_t s) {
  hipLaunchKernelGGL(k_dt, dim3((ninst + 63) / 64), dim3(64), 0, s, a, ninst);
}
// classify every tile (skipped by the marks, short path written, or
// listed), then a fixed grid walks the list with the full path
void launch_update(const StepArgs& a, hipStream_t s) {
  if (!a.dt_clears_count) launch_zero(a.tile_count, sizeof(int), s);
  TG_LAUNCH_TX(a.shape, k_update_classify, dim3(a.nblocks), s, a);
  TG_LAUNCH_TX(a.shape, k_update, dim3(a.ntiles < LIST_GRID ? a.ntiles : LIST_GRID), s, a);
}
void launch_adjoint(const AdjArgs& a, hipStream_t s) {
  launch_zero(a.tile_count, sizeof(int), s);
  TG_LAUNCH_TX(a.shape, k_adjoint_classify, dim3(a.nblocks), s, a);
  TG_LAUNCH_TX(a.shape, k_adjoint, dim3(a.ntiles < LIST_GRID ? a.ntiles : LIST_GRID), s, a);
}
static unsigned grid_for(long long n) {
  const long long b = (n + 255) / 256;
  return (unsigned)(b < 4096 ? (b < 1 ? 1 : b) : 4096);
}
void launch_zero(void* p, size_t bytes, hipStream_t s) {
  const long long n = (long long)(bytes / 4);
  hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(n)), dim3(256), 0, s, (uint32_t*)p, n);
}
void launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
  const long long n = (long long)(bytes / 4);
  hipLaunchKernelGGL(k_copy_u32, dim3(grid_for(n)), dim3(256), 0, s, (const uint32_t*)src,
                     (uint32_t*)dst, n);
}
template <class S>
void launch_box_max(const BoxChunk* ch, int nchunks, const long long* cells, const S* h,
                    double* rec, int max_steps, int step, hipStream_t s) {
  if (nchunks <= 0) return;
  hipLaunchKernelGGL(k_box_max<S>, dim3(nchunks), dim3(256), 0, s, ch, cells, h,
                     (unsigned long long*)rec, max_steps, step);
}
template void launch_box_max<double>(const BoxChunk*, int, const long long*, const double*,
                                     double*, int, int, hipStream_t);
template void launch_box_max<float>(const BoxChunk*, int, const long long*, const float*,
                                    double*, int, int, hipStream_t);
void launch_member_fill(const void* src, void* dst, int es, long long ncell, long long off,
                        long long count, uint32_t* mk, long long mw0, long long nmw,
                        hipStream_t s) {
  const long long w = es / 4;  // 32-bit words per value
  const long long n = count * w > nmw ? count * w : nmw;
  hipLaunchKernelGGL(k_member_fill, dim3(grid_for(n)), dim3(256), 0, s, (const uint32_t*)src,
                     (uint32_t*)dst, ncell * w, off * w, count * w, mk, mw0, nmw);
}
void launch_reset_red(RedDev* red, int ninst, hipStream_t s) {
  hipLaunchKernelGGL(k_reset_red, dim3((ninst + 63) / 64), dim3(64), 0, s, red, ninst);
}
void launch_stats(const InstDev* inst, const TileDev* tiles, int ntiles, Fields U,
                  const uint32_t* mk, double* part_mass, double* part_hmax, double tiny,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_stats, dim3(ntiles), dim3(BLOCK), 0, s, inst, tiles, U, mk, part_mass,
                     part_hmax, tiny);
}

void launch_copy_marking(const BlockDev* blocks, int nblocks, int shape, int es,
                         const void* const src[3], void* dst, long long ncell, uint32_t* mk_src,
                         uint32_t* mk_dst, hipStream_t s) {
  if (nblocks <= 0) return;
  if (es == 8) {
    const uint64_t* f[3] = {(const uint64_t*)src[0], (const uint64_t*)src[1],
                            (const uint64_t*)src[2]};
    switch (SHAPE_TX[shape]) {
      case 64: hipLaunchKernelGGL((k_copy_marking<64, uint64_t>), dim3(nblocks), dim3(BLOCK), 0,
                                  s, blocks, f[0], f[1], f[2], (uint64_t*)dst, ncell, mk_src,
                                  mk_dst); break;
      case 32: hipLaunchKernelGGL((k_copy_marking<32, uint64_t>), dim3(nblocks), dim3(BLOCK), 0,
                                  s, blocks, f[0], f[1], f[2], (uint64_t*)dst, ncell, mk_src,
                                  mk_dst); break;
      default: hipLaunchKernelGGL((k_copy_marking<16, uint64_t>), dim3(nblocks), dim3(BLOCK), 0,
                                  s, blocks, f[0], f[1], f[2], (uint64_t*)dst, ncell, mk_src,
                                  mk_dst); break;
    }
  } else {
    const uint32_t* f[3] = {(const uint32_t*)src[0], (const uint32_t*)src[1],
                            (const uint32_t*)src[2]};
    switch (SHAPE_TX[shape]) {
      case 64: hipLaunchKernelGGL((k_copy_marking<64, uint32_t>), dim3(nblocks), dim3(BLOCK), 0,
                                  s, blocks, f[0], f[1], f[2], (uint32_t*)dst, ncell, mk_src,
                                  mk_dst); break;
      case 32: hipLaunchKernelGGL((k_copy_marking<32, uint32_t>), dim3(nblocks), dim3(BLOCK), 0,
                                  s, blocks, f[0], f[1], f[2], (uint32_t*)dst, ncell, mk_src,
                                  mk_dst); break;
      default: hipLaunchKernelGGL((k_copy_marking<16, uint32_t>), dim3(nblocks), dim3(BLOCK), 0,
                                  s, blocks, f[0], f[1], f[2], (uint32_t*)dst, ncell, mk_src,
                                  mk_dst); break;
    }
  }
}

void launch_member_reduce(const double* sum_part, const double* max_part, const int* off,
                          const int* idx, int n, const double* time, const double* dt_rec,
                          long long stride, int col, double* out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_member_reduce, dim3(n), dim3(BLOCK), 0, s, sum_part, max_part, off, idx,
                     n, time, dt_rec, stride, col, out);
}

}  // namespace tg
