// Mixed-precision kernels.  Same tiles, tile table and algorithm as the
// FP64 kernels (tg_kernels.hip).  Template parameter S is the storage type
// of the state and adjoint:
//   S = double (TG_MIXED): FP32 increments applied to FP64 state/adjoint;
//                          switch arguments in FP64 (W = double)
//   S = float  (TG_FP32):  FP32 storage, switches in FP32 (W = float)
// Static coefficients and kact are FP32 in both; the CFL reduction, dt,
// time, mass and the bed-friction gradient sums are FP64.
// Nothing here is held to the reference's operand order (that is the FP64
// kernels' job): let a*b+c contract to fused multiply-adds, also in the
// included FP32 math.
#pragma clang fp contract(fast)

#include <type_traits>

#include "tg_kernels.h"
#include "tg_math32.h"

namespace tg {

using namespace m32;

namespace {

template <class S>
using Wt = typename std::conditional<std::is_same<S, double>::value, double, float>::type;

__device__ __forceinline__ unsigned long long dbits(double v) {
  return (unsigned long long)__double_as_longlong(v);
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
  bool act, hasE, hasW, hasN, hasS;
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

// Tile t is known to hold +0.0 in all three fields of a marked buffer: its
// block's word is MARK_FULL, or its own byte is 1 and the bytes follow this
// tile shape (bytes != 0).
template <int TX>
__device__ __forceinline__ bool tile_marked_zero(const uint32_t* mk, const TileDev& t,
                                                 int bytes) {
  const uint32_t w = mk[t.mw];
  return w == MARK_FULL || (bytes && ((w >> (8 * mark_byte<TX>(t))) & 0xffu) == 1u);
}
template <class S>
__device__ __forceinline__ bool bit_zero(S v) {
  if constexpr (std::is_same<S, double>::value) return __double_as_longlong(v) == 0ll;
  else return __float_as_uint(v) == 0u;
}
// the tile's byte of the output marks (thread 0 stores it)
template <int TX>
__device__ __forceinline__ void tile_mark_store(uint32_t* mk, const TileDev& t, bool zero) {
  if (threadIdx.x == 0) ((unsigned char*)(mk + t.mw))[mark_byte<TX>(t)] = zero ? 1 : 0;
}

// The tile next to t on side s (0 W, 1 E, 2 S, 3 N) holds +0.0 per the
// marks, or there is none (t lies on the member's edge).
template <int TX>
__device__ __forceinline__ bool side_tile_zero(const uint32_t* mk, const TileDev& t,
                                               const InstDev& I, int s, int bytes) {
  constexpr int TY = BLOCK / TX;
  TileDev n = t;
  int w;
  if (s == 0) {
    if (t.x0 == 0) return true;
    n.x0 = t.x0 - TX;
    w = (t.x0 % MARK_W == 0) ? t.mwW : t.mw;
  } else if (s == 1) {
    if (t.x0 + TX >= I.nx) return true;
    n.x0 = t.x0 + TX;
    w = (n.x0 % MARK_W == 0) ? t.mwE : t.mw;
  } else if (s == 2) {
    if (t.y0 == 0) return true;
    n.y0 = t.y0 - TY;
    w = (t.y0 % MARK_H == 0) ? t.mwS : t.mw;
  } else {
    if (t.y0 + TY >= I.ny) return true;
    n.y0 = t.y0 + TY;
    w = (n.y0 % MARK_H == 0) ? t.mwN : t.mw;
  }
  if (w < 0) return false;  // (a neighbour inside the member has a block)
  n.mw = w;
  return tile_marked_zero<TX>(mk, n, bytes);
}

// ---------------------------------------------------------------- forward --

template <class S, int TX>
__global__ __launch_bounds__(BLOCK) void k32_coef(StepArgsM<S> a, int reduce) {
  using W = Wt<S>;
  const TileDev t = a.tiles[blockIdx.x];
  const InstDev I = a.inst[t.inst];
  // a tile of U marked +0.0: h = +0.0 < tiny everywhere -- no kact read
  // (the update and the adjoint take kact only above tiny), nothing for the
  // CFL reduction
  if (a.mk_in && tile_marked_zero<TX>(a.mk_in, t, a.zbytes)) return;
  const Cell q = cell_of<TX>(t, I);
  const W tiny = (W)a.tiny, gamma = (W)a.gamma;
  double minev = HUGE_VAL, hmax = 0.;
  if (q.act) {
    const long long c = q.c, nx = I.nx;
    const W h = (W)a.U.h[c], hu = (W)a.U.hu[c], hv = (W)a.U.hv[c];
    const W ix = (W)I.inv_dx, iy = (W)I.inv_dy;
    float k = (float)I.eps;
    if (h > tiny) {  // (kact reads the slopes only above tiny)
      W hx = 0, hux = 0, hy = 0, hvy = 0;
      if (q.hasE && q.hasW && q.hasN && q.hasS) {
        hx = slope<W>((W)a.U.h[c + 1], h, (W)a.U.h[c - 1], ix, gamma, tiny);
        hux = slope<W>((W)a.U.hu[c + 1], hu, (W)a.U.hu[c - 1], ix, gamma, tiny);
        hy = slope<W>((W)a.U.h[c + nx], h, (W)a.U.h[c - nx], iy, gamma, tiny);
        hvy = slope<W>((W)a.U.hv[c + nx], hv, (W)a.U.hv[c - nx], iy, gamma, tiny);
      }
      const W ih = inv_h(h);
      const W u = vel(hu, h, ih), v = vel(hv, h, ih);
      k = kact<W>(h, u, v, ih, hx, hux, hy, hvy, (float)I.kc1,
                  (float)mat_of(I, a.mat, a.mtab, c).kroot, (float)I.eps, tiny);
      if (reduce) {
        const float c2 = fsqrt(k * a.S32.gz[c] * (float)h);
        const double ev = (double)fmax(fabs(u), fabs(v)) + (double)c2;
        if (ev > 1000000000.) atomicOr(a.err, 2);
        minev = ((I.dx > I.dy) ? I.dy : I.dx) / ev;
        hmax = (double)h;
      }
    }
    a.kact[c] = k;
  }
  if (reduce) {
    __shared__ double smin[BLOCK / 64], smax[BLOCK / 64];
    minev = wave_min(minev);
    hmax = wave_max(hmax);
    if (threadIdx.x % 64 == 0) { smin[threadIdx.x / 64] = minev; smax[threadIdx.x / 64] = hmax; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < BLOCK / 64; k++) {
        minev = (smin[k] < minev) ? smin[k] : minev;
        hmax = (smax[k] > hmax) ? smax[k] : hmax;
      }
      if (minev < HUGE_VAL) atomicMin(&a.red[t.inst].min_ev, dbits(minev));
      if (hmax > 0.) atomicMax(&a.red[t.inst].hmax, dbits(hmax));
    }
  }
}

// Records of a tile and its one-cell frame, staged once per workgroup: the
// state in the switch type (slopes, the wet test, sources) and the inputs of
// the physical flux (hfv_at's normal velocities, kact, g_z); then every face
// flux of the tile once (a face is shared by two cells).
template <class W, int TX>
struct Lds32 {
  static constexpr int TY = BLOCK / TX, RX = TX + 2, RXY = (TX + 2) * (TY + 2);
  W h[RXY], hu[RXY], hv[RXY];
  float ux[RXY], uy[RXY], k[RXY], g[RXY];
  int wet[RXY];
  float fx[3][TY][TX + 1], fy[3][TY + 1][TX];
};

template <class W, int TX>
__device__ __forceinline__ Hfv rec_hfv32(const Lds32<W, TX>& L, int dir, int r) {
  return hfv(dir, (float)L.h[r], (float)L.hu[r], (float)L.hv[r], dir ? L.uy[r] : L.ux[r], L.k[r],
             L.g[r], L.wet[r] != 0);
}

// One workgroup per tile.  A tile whose cells and frame are all below tiny
// has every face flux zero and no sources: its cells take the update's own
// arithmetic with zero fluxes, without the flux and slope work.
template <class S, int TX>
__global__ __launch_bounds__(BLOCK) void k32_update(StepArgsM<S> a) {
  using W = Wt<S>;
  using L32 = Lds32<W, TX>;
  constexpr int TY = L32::TY, RX = L32::RX, RXY = L32::RXY;
  __shared__ L32 L;
  const TileDev t = a.tiles[blockIdx.x];
  const InstDev I = a.inst[t.inst];
  const double dt = a.dt_rec[(long long)t.inst * a.max_steps + a.step];
  // the tile and its four side tiles of U marked +0.0: zero fluxes, no
  // sources (h < tiny), so U^{s+1} = +0.0 + (-dtdx * 0 - dtdy * 0) = +0.0 for
  // a finite dt; nothing to do when the output tile is marked +0.0 already
  if (a.mk_in && a.mk_out && isfinite(dt) && tile_marked_zero<TX>(a.mk_out, t, a.zbytes) &&
      tile_marked_zero<TX>(a.mk_in, t, a.zbytes) && side_tile_zero<TX>(a.mk_in, t, I, 0, a.zbytes) &&
      side_tile_zero<TX>(a.mk_in, t, I, 1, a.zbytes) &&
      side_tile_zero<TX>(a.mk_in, t, I, 2, a.zbytes) &&
      side_tile_zero<TX>(a.mk_in, t, I, 3, a.zbytes))
    return;
  const Cell q = cell_of<TX>(t, I);
  const long long nx = I.nx;
  const W tiny = (W)a.tiny, gamma = (W)a.gamma;
  // phase 0: records of the tile and its frame (cells of the member only)
  bool any_wet = false;
  for (int r = (int)threadIdx.x; r < RXY; r += BLOCK) {
    const int gi = t.x0 - 1 + r % RX, gj = t.y0 - 1 + r / RX;
    if (gi >= 0 && gi < I.nx && gj >= 0 && gj < I.ny) {
      const long long c = I.off + (long long)gj * nx + gi;
      const W h = (W)a.U.h[c], hu = (W)a.U.hu[c], hv = (W)a.U.hv[c];
      const bool wet = !(h < tiny);
      L.h[r] = h; L.hu[r] = hu; L.hv[r] = hv;
      L.wet[r] = wet ? 1 : 0;
      float ux = 0.f, uy = 0.f, k = 0.f, g = 0.f;
      if (wet) {
        ux = (float)vel(hu, h, inv_h(h));
        uy = (float)vel(hv, h, inv_h(h));
        k = a.kact[c];
        g = a.S32.gz[c];
        any_wet = true;
      }
      L.ux[r] = ux; L.uy[r] = uy; L.k[r] = k; L.g[r] = g;
    }
  }
  const bool tile_wet = __syncthreads_or(any_wet);
  if (tile_wet) {
    // phase 1: faces with a cell of the member on both sides
    for (int f = (int)threadIdx.x; f < (TX + 1) * TY; f += BLOCK) {
      const int fi = f % (TX + 1), fj = f / (TX + 1);
      const int gi = t.x0 + fi, gj = t.y0 + fj;
      if (gi > 0 && gi < I.nx && gj < I.ny) {
        const int r = (fj + 1) * RX + fi + 1;
        float fl[3];
        riemann(rec_hfv32<W, TX>(L, 0, r - 1), rec_hfv32<W, TX>(L, 0, r), fl);
        for (int v = 0; v < 3; v++) L.fx[v][fj][fi] = fl[v];
      }
    }
    for (int f = (int)threadIdx.x; f < TX * (TY + 1); f += BLOCK) {
      const int fi = f % TX, fj = f / TX;
      const int gi = t.x0 + fi, gj = t.y0 + fj;
      if (gj > 0 && gj < I.ny && gi < I.nx) {
        const int r = (fj + 1) * RX + fi + 1;
        float fl[3];
        riemann(rec_hfv32<W, TX>(L, 1, r - RX), rec_hfv32<W, TX>(L, 1, r), fl);
        for (int v = 0; v < 3; v++) L.fy[v][fj][fi] = fl[v];
      }
    }
    __syncthreads();
  }
  bool zero_out = true;  // this cell's U^{s+1} is +0.0 (or no cell)
  if (q.act) {
    const long long c = q.c;
    const int tx = (int)(threadIdx.x % TX), ty = (int)(threadIdx.x / TX);
    const int rs = (ty + 1) * RX + tx + 1;
    const float dtf = (float)dt, dtdx = (float)(dt * I.inv_dx), dtdy = (float)(dt * I.inv_dy);
    const S hs = a.U.h[c], hus = a.U.hu[c], hvs = a.U.hv[c];
    const W h = L.h[rs], hu = L.hu[rs], hv = L.hv[rs];
    float fE[3] = {0.f, 0.f, 0.f}, fW[3] = {0.f, 0.f, 0.f}, fN[3] = {0.f, 0.f, 0.f},
          fS[3] = {0.f, 0.f, 0.f};
    W hx = 0, hvx = 0, hy = 0, huy = 0;
    if (tile_wet) {
      const W ix = (W)I.inv_dx, iy = (W)I.inv_dy;
      if (q.hasE && q.hasW && q.hasN && q.hasS) {
        hx = slope<W>(L.h[rs + 1], h, L.h[rs - 1], ix, gamma, tiny);
        hvx = slope<W>(L.hv[rs + 1], hv, L.hv[rs - 1], ix, gamma, tiny);
        hy = slope<W>(L.h[rs + RX], h, L.h[rs - RX], iy, gamma, tiny);
        huy = slope<W>(L.hu[rs + RX], hu, L.hu[rs - RX], iy, gamma, tiny);
      }
#pragma unroll
      for (int v = 0; v < 3; v++) {
        if (q.hasE) fE[v] = L.fx[v][ty][tx + 1];
        if (q.hasW) fW[v] = L.fx[v][ty][tx];
        if (q.hasN) fN[v] = L.fy[v][ty + 1][tx];
        if (q.hasS) fS[v] = L.fy[v][ty][tx];
      }
#pragma unroll
      for (int v = 0; v < 3; v++) {  // outflow copies
        if (!q.hasE) fE[v] = fW[v];
        if (!q.hasW) fW[v] = fE[v];
        if (!q.hasN) fN[v] = fS[v];
        if (!q.hasS) fS[v] = fN[v];
      }
    }
    // increments in FP32, applied to the stored state type S
    float inc[3];
#pragma unroll
    for (int v = 0; v < 3; v++) inc[v] = -dtdx * (fE[v] - fW[v]) - dtdy * (fN[v] - fS[v]);
    if (h > tiny) {
      const W ih = inv_h(h);
      const W u = vel(hu, h, ih), v = vel(hv, h, ih);
      const W ft = (W)a.ft;
      const float sgn0 = (float)tiny_sgn<W>(ih * (huy - u * hy), ft);
      const float sgn2 = (float)tiny_sgn<W>(u, ft);
      const float sgn1 = (float)tiny_sgn<W>(ih * (hvx - v * hx), ft);
      const float sgn3 = (float)tiny_sgn<W>(v, ft);
      const float k = a.kact[c], g2 = a.S32.gz[c];
      const float hf = (float)h, uf = (float)u, vf = (float)v, huf = (float)hu, hvf = (float)hv;
      const float sint = (float)I.sin_int, tanb = (float)mat_of(I, a.mat, a.mtab, c).tan_bed;
      float s1 = a.S32.gx[c] * hf;
      float s2 = sgn0 * hf * k * (g2 * (float)hy + a.S32.dgy[c] * hf) * sint;
      float s3 = sgn2 * fmaxf(g2 * hf + uf * huf * a.S32.kx[c], 0.f) * tanb;
      inc[1] += dtf * (s1 - s2 - s3);
      s1 = a.S32.gy[c] * hf;
      s2 = sgn1 * hf * k * (g2 * (float)hx + a.S32.dgx[c] * hf) * sint;
      s3 = sgn3 * fmaxf(g2 * hf + vf * hvf * a.S32.ky[c], 0.f) * tanb;
      inc[2] += dtf * (s1 - s2 - s3);
    }
    S out0 = hs + (S)inc[0];
    if (out0 < (S)0) out0 = (S)0;
    const S out1 = hus + (S)inc[1], out2 = hvs + (S)inc[2];
    if (!isfinite((double)out0) || !isfinite((double)out1) || !isfinite((double)out2))
      atomicOr(a.err, 1);
    a.h_out[c] = out0;
    a.hu_out[c] = out1;
    a.hv_out[c] = out2;
    zero_out = bit_zero(out0) && bit_zero(out1) && bit_zero(out2);
  }
  if (a.mk_out) tile_mark_store<TX>(a.mk_out, t, __syncthreads_and(zero_out));
}

__global__ __launch_bounds__(BLOCK) void k32_stats(const InstDev* inst, const TileDev* tiles,
                                                   FieldsT<float> U, double* part_mass,
                                                   double* part_hmax, double tiny) {
  const TileDev t = tiles[blockIdx.x];
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

// ---------------------------------------------------------------- adjoint --
//
// The Jacobian blocks are formed in arithmetic type F: float for the bulk;
// double for cells next to a thin layer (0 < h < THIN in the cell or a
// neighbour).  At wetting fronts the reverse recursion has growing modes
// (1/h terms of the thin cell's momentum Jacobian) that amplify whatever
// rounding seeds them; FP32 seeds them some 1e8 times larger than FP64.

template <class F>
struct M3T {
  F a[3][3];
};
template <class F>
__device__ __forceinline__ M3T<F> zero3() {
  M3T<F> m;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) m.a[i][j] = F(0);
  return m;
}

// Static coefficient k (0 gx, 1 gy, 2 gz, 3 dgx, 4 dgy, 5 kx, 6 ky) of cell c
// in arithmetic type F: the FP64 copy when F is double and one exists (the
// thin-layer pass of TG_FWD64_ADJ32), else the FP32 array.
template <class F, class S>
__device__ __forceinline__ F stat(const AdjArgsM<S>& a, int k, long long c) {
  if constexpr (sizeof(F) == 8) {
    if (a.s64) return a.s64[(long long)k * a.ncell + c];
  }
  const float* p[7] = {a.S32.gx, a.S32.gy, a.S32.gz, a.S32.dgx, a.S32.dgy, a.S32.kx, a.S32.ky};
  return (F)p[k][c];
}

// What the Jacobian blocks need about one cell at U^s.  Switch decisions
// (slopes, hslope branch, kact sign, sgn, stop, wet, thick) are taken in W
// before the values are narrowed to F.
template <class F>
struct DualInfoT {
  F h, hu, hv, g2, ih, u, v, k, dk;
  F hx, hy;           // own h slopes
  F hsx[3], hsy[3];   // hslope_sens
  F sgn[4];
  F lam[3];           // lambda^{s+1} of the cell
  int stop[2];
  int wet;            // !(h < tiny)
  int thick;          // h > tiny
  int nz;             // h != 0
};

template <class F>
struct DualHfvT {
  HfvT<F> b;
  M3T<F> fj, sj;
};
template <class F>
__device__ __forceinline__ DualHfvT<F> dual_hfv(int dir, const DualInfoT<F>& d) {
  DualHfvT<F> r;
  const F u = d.u, v = d.v, ih = d.ih, h = d.h;
  r.b = hfv<F>(dir, h, d.hu, d.hv, dir ? v : u, d.k, d.g2, d.wet);
  r.fj = zero3<F>();
  r.sj = zero3<F>();
  if (!d.wet) return r;
  const F a2 = d.k * h * d.g2, a = fsqrt(a2);
  if (dir == 0) {
    r.fj.a[0][1] = F(1);
    r.fj.a[1][0] = a2 - u * u; r.fj.a[1][1] = F(2) * u;
    r.fj.a[2][0] = -u * v; r.fj.a[2][1] = v; r.fj.a[2][2] = u;
    r.sj.a[0][0] = (F(-0.5) * a - u) * ih; r.sj.a[0][1] = ih;
    r.sj.a[1][0] = -u * ih; r.sj.a[1][1] = ih;
    r.sj.a[2][0] = (F(0.5) * a - u) * ih; r.sj.a[2][1] = ih;
  } else {
    r.fj.a[0][2] = F(1);
    r.fj.a[1][0] = -v * u; r.fj.a[1][1] = v; r.fj.a[1][2] = u;
    r.fj.a[2][0] = a2 - v * v; r.fj.a[2][2] = F(2) * v;
    r.sj.a[0][0] = (F(-0.5) * a - v) * ih; r.sj.a[0][2] = ih;
    r.sj.a[1][0] = -v * ih; r.sj.a[1][2] = ih;
    r.sj.a[2][0] = (F(0.5) * a - v) * ih; r.sj.a[2][2] = ih;
  }
  return r;
}

// d(HLL face flux)/d(U of one side), branch logic of dual_riemannflux
// (dualmesh.C:1628-1843)
template <class F>
__device__ __forceinline__ M3T<F> dual_riemann_side(const DualHfvT<F>& l, const DualHfvT<F>& r,
                                                    bool want_left) {
  M3T<F> j = zero3<F>();
  const SpeedsT<F> s = speeds(l.b, r.b);
  if (s.mode == 0) return j;
  if (s.mode == 1) return want_left ? l.fj : j;
  if (s.mode == 2) return want_left ? j : r.fj;
  F dsl[3], dsr[3];
  const DualHfvT<F>& me = want_left ? l : r;
  const bool ldry = l.b.u[0] == F(0), rdry = r.b.u[0] == F(0);
#pragma unroll
  for (int i = 0; i < 3; i++) { dsl[i] = F(0); dsr[i] = F(0); }
  if (ldry || rdry) {
    const bool mine_wet = want_left ? !ldry : !rdry;
    if (mine_wet) {
#pragma unroll
      for (int i = 0; i < 3; i++) {
        dsl[i] = (s.sl == F(0)) ? F(0) : F(2) * me.sj.a[0][i] - me.sj.a[1][i];
        dsr[i] = (s.sr == F(0)) ? F(0) : F(2) * me.sj.a[2][i] - me.sj.a[1][i];
      }
    }
  } else {
    const bool sl_left = s.sl == l.b.w[0], sr_left = s.sr == l.b.w[2];
    const bool sl_mine = (s.sl != F(0)) && (want_left ? sl_left : !sl_left);
    const bool sr_mine = (s.sr != F(0)) && (want_left ? sr_left : !sr_left);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      dsl[i] = sl_mine ? me.sj.a[0][i] : F(0);
      dsr[i] = sr_mine ? me.sj.a[2][i] : F(0);
    }
  }
  const F inv = rcp(s.sr - s.sl);
  F dfdsr[3], dfdsl[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const F vd = r.b.u[i] - l.b.u[i];
    const F flux = (s.sr * l.b.f[i] - s.sl * r.b.f[i] + s.sl * s.sr * vd) * inv;
    dfdsr[i] = (l.b.f[i] + s.sl * vd - flux) * inv;
    dfdsl[i] = (-r.b.f[i] + s.sr * vd + flux) * inv;
  }
  const F dF = want_left ? s.sr * inv : -s.sl * inv;
  const F jv = want_left ? -s.sr * s.sl * inv : s.sr * s.sl * inv;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++)
      j.a[i][k] = dfdsr[i] * dsr[k] + dfdsl[i] * dsl[k] + dF * me.fj.a[i][k] + ((i == k) ? jv : F(0));
  return j;
}

template <class S, class F>
__device__ __forceinline__ DualInfoT<F> dual_info(const AdjArgsM<S>& a, const InstDev& I, int i,
                                                  int j, bool want_dk) {
  using W = Wt<S>;
  DualInfoT<F> r;
  const W tiny = (W)a.tiny, gamma = (W)a.gamma;
  const long long nx = I.nx, c = I.off + (long long)j * nx + i;
  const W h = (W)a.U.h[c], hu = (W)a.U.hu[c], hv = (W)a.U.hv[c];
  W d[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; k++) r.hsx[k] = r.hsy[k] = F(0);
  if (i + 1 < I.nx && i > 0 && j + 1 < I.ny && j > 0) {
    const W ix = (W)I.inv_dx, iy = (W)I.inv_dy;
    Slope<W> s = slope_sens<W>((W)a.U.h[c + 1], h, (W)a.U.h[c - 1], ix, gamma, tiny);
    d[0] = s.d; r.hsx[0] = (F)s.s0; r.hsx[1] = (F)s.s1; r.hsx[2] = (F)s.s2;
    d[1] = slope<W>((W)a.U.hu[c + 1], hu, (W)a.U.hu[c - 1], ix, gamma, tiny);
    d[2] = slope<W>((W)a.U.hv[c + 1], hv, (W)a.U.hv[c - 1], ix, gamma, tiny);
    s = slope_sens<W>((W)a.U.h[c + nx], h, (W)a.U.h[c - nx], iy, gamma, tiny);
    d[3] = s.d; r.hsy[0] = (F)s.s0; r.hsy[1] = (F)s.s1; r.hsy[2] = (F)s.s2;
    d[4] = slope<W>((W)a.U.hu[c + nx], hu, (W)a.U.hu[c - nx], iy, gamma, tiny);
    d[5] = slope<W>((W)a.U.hv[c + nx], hv, (W)a.U.hv[c - nx], iy, gamma, tiny);
  }
  const W ih = inv_h(h);
  const W u = (h > W(0)) ? vel(hu, h, ih) : W(0), v = (h > W(0)) ? vel(hv, h, ih) : W(0);
  r.h = (F)h; r.hu = (F)hu; r.hv = (F)hv; r.g2 = stat<F>(a, 2, c);
  r.lam[0] = (F)a.lam_next.h[c]; r.lam[1] = (F)a.lam_next.hu[c]; r.lam[2] = (F)a.lam_next.hv[c];
  r.ih = (F)ih; r.u = (F)u; r.v = (F)v;
  r.wet = !(h < tiny);
  r.thick = h > tiny;
  r.nz = h != W(0);
  r.hx = (F)d[0];
  r.hy = (F)d[3];
  const MatDev M = mat_of(I, a.mat, a.mtab, c);
  r.k = kact<W, F>(h, u, v, ih, d[0], d[1], d[3], d[5], (F)I.kc1, (F)M.kroot, (F)I.eps, tiny);
  r.dk = want_dk ? dkact_dphi<W, F>(h, u, v, ih, d[0], d[1], d[3], d[5], (F)I.kc1, (F)M.kroot,
                                    (F)M.dkroot, (F)I.eps, tiny)
                 : F(0);
  r.sgn[0] = r.sgn[1] = r.sgn[2] = r.sgn[3] = F(0);
  r.stop[0] = r.stop[1] = 0;
  if (r.thick) {
    const W ft = (W)a.ft;
    r.sgn[0] = (F)tiny_sgn<W>(ih * (d[4] - u * d[3]), ft);
    r.sgn[2] = (F)tiny_sgn<W>(u, ft);
    r.sgn[1] = (F)tiny_sgn<W>(ih * (d[2] - v * d[0]), ft);
    r.sgn[3] = (F)tiny_sgn<W>(v, ft);
    // stop: the Coulomb term's max(.,0) clamps to zero (or tan(phi_bed) = 0)
    const W g2 = (W)stat<F>(a, 2, c), tanb = (W)M.tan_bed;
    r.stop[0] = r.sgn[2] != F(0) && (!(g2 * h + u * hu * (W)stat<F>(a, 5, c) > W(0)) || tanb == W(0));
    r.stop[1] = r.sgn[3] != F(0) && (!(g2 * h + v * hv * (W)stat<F>(a, 6, c) > W(0)) || tanb == W(0));
  }
  return r;
}

// LDS record slots: the largest tile plus halo over the built tile widths
constexpr int RN = (64 + 2) * (BLOCK / 64 + 2);
struct DualLds {
  float h[RN], hu[RN], hv[RN], g2[RN], ih[RN], u[RN], v[RN], k[RN], dk[RN];
  float hsx1[RN], hsx2[RN], hsy1[RN], hsy2[RN];
  float lh[RN], lhu[RN], lhv[RN];  // lambda^{s+1}
  int flags[RN];  // sgn0+1 | (sgn1+1)<<2 | stop0<<4 | stop1<<5 | wet<<6 | thick<<7 | nz<<8
};
__device__ __forceinline__ void lds_put(DualLds& L, int r, const DualInfoT<float>& d) {
  L.h[r] = d.h; L.hu[r] = d.hu; L.hv[r] = d.hv; L.g2[r] = d.g2; L.ih[r] = d.ih;
  L.u[r] = d.u; L.v[r] = d.v; L.k[r] = d.k; L.dk[r] = d.dk;
  L.hsx1[r] = d.hsx[1]; L.hsx2[r] = d.hsx[2]; L.hsy1[r] = d.hsy[1]; L.hsy2[r] = d.hsy[2];
  L.lh[r] = d.lam[0]; L.lhu[r] = d.lam[1]; L.lhv[r] = d.lam[2];
  L.flags[r] = (int)(d.sgn[0] + 1.f) | ((int)(d.sgn[1] + 1.f) << 2) | (d.stop[0] << 4) |
               (d.stop[1] << 5) | (d.wet << 6) | (d.thick << 7) | (d.nz << 8);
}
__device__ __forceinline__ DualInfoT<float> lds_get(const DualLds& L, int r) {
  DualInfoT<float> d;
  d.h = L.h[r]; d.hu = L.hu[r]; d.hv = L.hv[r]; d.g2 = L.g2[r]; d.ih = L.ih[r];
  d.u = L.u[r]; d.v = L.v[r]; d.k = L.k[r]; d.dk = L.dk[r];
  d.hsx[0] = d.hsy[0] = 0.f;
  d.hsx[1] = L.hsx1[r]; d.hsx[2] = L.hsx2[r]; d.hsy[1] = L.hsy1[r]; d.hsy[2] = L.hsy2[r];
  d.lam[0] = L.lh[r]; d.lam[1] = L.lhu[r]; d.lam[2] = L.lhv[r];
  const int f = L.flags[r];
  d.sgn[0] = (float)((f & 3) - 1);
  d.sgn[1] = (float)(((f >> 2) & 3) - 1);
  d.sgn[2] = d.sgn[3] = 0.f;
  d.stop[0] = (f >> 4) & 1;
  d.stop[1] = (f >> 5) & 1;
  d.wet = (f >> 6) & 1;
  d.thick = (f >> 7) & 1;
  d.nz = (f >> 8) & 1;
  d.hx = d.hy = 0.f;
  return d;
}

template <class F>
__device__ __forceinline__ void accum_T(F adj[3], const M3T<F>& J, const F lam[3]) {
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int l = 0; l < 3; l++) adj[k] += lam[l] * J.a[l][k];
}

// neighbour-block source part: slope sensitivity of the internal friction
template <class F>
__device__ __forceinline__ void nbr_block(M3T<F>& J, const DualInfoT<F>& n, F sin_int, F dt,
                                          F dh0, F dh1) {
  if (n.thick) {
    const F alpha = sin_int * n.sgn[0] * n.k, betta = sin_int * n.sgn[1] * n.k;
    if (!n.stop[0]) J.a[1][0] += dt * n.h * alpha * n.g2 * dh1;
    if (!n.stop[1]) J.a[2][0] += dt * n.h * betta * n.g2 * dh0;
  }
}

// d(face flux)/dk of one side
template <class F>
__device__ __forceinline__ void dhfv_dk(int dir, const DualInfoT<F>& d, F dF[3], F dW[3]) {
  dF[0] = dF[1] = dF[2] = dW[0] = dW[1] = dW[2] = F(0);
  if (!d.wet) return;
  const F a = fsqrt(d.k * d.h * d.g2);
  const F da = (a > F(0)) ? F(0.5) * d.h * d.g2 * rcp(a) : F(0);
  dF[dir ? 2 : 1] = F(0.5) * d.h * d.g2 * d.h;
  dW[0] = -da;
  dW[2] = da;
}

template <class F>
__device__ __forceinline__ void driemann(const HfvT<F>& l, const HfvT<F>& r, const F dFl[3],
                                         const F dWl[3], const F dFr[3], const F dWr[3],
                                         F out[3]) {
  out[0] = out[1] = out[2] = F(0);
  const SpeedsT<F> s = speeds(l, r);
  if (s.mode == 0) return;
  if (s.mode == 1) { for (int v = 0; v < 3; v++) out[v] = dFl[v]; return; }
  if (s.mode == 2) { for (int v = 0; v < 3; v++) out[v] = dFr[v]; return; }
  F dsl, dsr;
  if (l.u[0] == F(0)) {
    dsl = (s.sl < F(0)) ? F(2) * dWr[0] - dWr[1] : F(0);
    dsr = (s.sr > F(0)) ? F(2) * dWr[2] - dWr[1] : F(0);
  } else if (r.u[0] == F(0)) {
    dsl = (s.sl < F(0)) ? F(2) * dWl[0] - dWl[1] : F(0);
    dsr = (s.sr > F(0)) ? F(2) * dWl[2] - dWl[1] : F(0);
  } else {
    dsl = (s.sl < F(0)) ? ((l.w[0] > r.w[0]) ? dWr[0] : dWl[0]) : F(0);
    dsr = (s.sr > F(0)) ? ((l.w[2] > r.w[2]) ? dWl[2] : dWr[2]) : F(0);
  }
  const F inv = rcp(s.sr - s.sl);
#pragma unroll
  for (int v = 0; v < 3; v++) {
    const F vd = r.u[v] - l.u[v];
    const F f = (s.sr * l.f[v] - s.sl * r.f[v] + s.sl * s.sr * vd) * inv;
    const F dfdsr = (l.f[v] + s.sl * vd - f) * inv, dfdsl = (-r.f[v] + s.sr * vd + f) * inv;
    out[v] = dfdsr * dsr + dfdsl * dsl + s.sr * inv * dFl[v] - s.sl * inv * dFr[v];
  }
}

// face flux and its derivative along d kact/d phi_bed (both sides)
template <class F>
__device__ __forceinline__ void face_dk(int side, const DualInfoT<F>& L, const DualInfoT<F>& R,
                                        F flux[3], F dflux[3]) {
  const HfvT<F> hl = hfv<F>(side, L.h, L.hu, L.hv, side ? L.v : L.u, L.k, L.g2, L.wet);
  const HfvT<F> hr = hfv<F>(side, R.h, R.hu, R.hv, side ? R.v : R.u, R.k, R.g2, R.wet);
  riemann(hl, hr, flux);
  F dFl[3], dWl[3], dFr[3], dWr[3];
  dhfv_dk(side, L, dFl, dWl);
  dhfv_dk(side, R, dFr, dWr);
#pragma unroll
  for (int v = 0; v < 3; v++) {
    dFl[v] *= L.dk; dWl[v] *= L.dk; dFr[v] *= R.dk; dWr[v] *= R.dk;
  }
  driemann(hl, hr, dFl, dWl, dFr, dWr, dflux);
}

// lambda^{s+1} . dF/d phi_bed of one cell at frozen dt
template <class S, class F>
__device__ __forceinline__ F dphi_term(const AdjArgsM<S>& a, const InstDev& I, long long c,
                       const DualInfoT<F>& C, const DualInfoT<F>& E, const DualInfoT<F>& N,
                       const DualInfoT<F>& Wc, const DualInfoT<F>& Sc, const bool has[5], F dt,
                       const F lam[3]) {
  F fE[3] = {0, 0, 0}, fN[3] = {0, 0, 0}, fW[3] = {0, 0, 0}, fS[3] = {0, 0, 0};
  F dE[3] = {0, 0, 0}, dN[3] = {0, 0, 0}, dW[3] = {0, 0, 0}, dS[3] = {0, 0, 0};
  if (has[1]) face_dk(0, C, E, fE, dE);
  if (has[3]) face_dk(0, Wc, C, fW, dW);
  if (has[2]) face_dk(1, C, N, fN, dN);
  if (has[4]) face_dk(1, Sc, C, fS, dS);
  // boundary faces copy the opposite face
#pragma unroll
  for (int v = 0; v < 3; v++) {
    if (!has[1]) { fE[v] = fW[v]; dE[v] = dW[v]; }
    if (!has[3]) { fW[v] = fE[v]; dW[v] = dE[v]; }
    if (!has[2]) { fN[v] = fS[v]; dN[v] = dS[v]; }
    if (!has[4]) { fS[v] = fN[v]; dS[v] = dN[v]; }
  }
  const F dtdx = (F)(dt * I.inv_dx), dtdy = (F)(dt * I.inv_dy);
  F out[3];
#pragma unroll
  for (int v = 0; v < 3; v++) out[v] = -dtdx * (dE[v] - dW[v]) - dtdy * (dN[v] - dS[v]);
  if (C.h - dtdx * (fE[0] - fW[0]) - dtdy * (fN[0] - fS[0]) < F(0)) out[0] = F(0);
  if (C.thick) {
    const double tbc = mat_of(I, a.mat, a.mtab, c).tan_bed;
    const F sec2 = (F)(1. + tbc * tbc), sint = (F)I.sin_int;
    const F g2 = C.g2;
    F ds2x = C.sgn[0] * C.h * C.dk * (g2 * C.hy + stat<F>(a, 4, c) * C.h) * sint;
    F ds3x = C.sgn[2] * fmax(g2 * C.h + C.u * C.hu * stat<F>(a, 5, c), F(0)) * sec2;
    F ds2y = C.sgn[1] * C.h * C.dk * (g2 * C.hx + stat<F>(a, 3, c) * C.h) * sint;
    F ds3y = C.sgn[3] * fmax(g2 * C.h + C.v * C.hv * stat<F>(a, 6, c), F(0)) * sec2;
    out[1] -= dt * (ds2x + ds3x);
    out[2] -= dt * (ds2y + ds3y);
  }
  return lam[0] * out[0] + lam[1] * out[1] + lam[2] * out[2];
}

// The gather of one cell in arithmetic type F: adj = sum of the transposed
// Jacobian blocks (the self block without its identity) applied to
// lambda^{s+1}; nbr(s) returns the record of neighbour slot s (1 E, 2 N,
// 3 W, 4 S).  Returns the cell's dJ/dphi_bed term (0 unless want_dphi).
template <class S, class F, class NBR>
__device__ __forceinline__ double gather(const AdjArgsM<S>& a, const InstDev& I, const Cell& q,
                                         const DualInfoT<F>& self, NBR&& nbr, double dtd,
                                         F adj[3], bool want_dphi) {
  const long long c = q.c, nx = I.nx;
  const F dt = (F)dtd, dtdx = (F)(dtd * I.inv_dx), dtdy = (F)(dtd * I.inv_dy);
  const bool has[5] = {true, q.hasE, q.hasN, q.hasW, q.hasS};
  adj[0] = adj[1] = adj[2] = F(0);
  if (self.nz) {
    M3T<F> JE = zero3<F>(), JN = zero3<F>(), JW = zero3<F>(), JS = zero3<F>();
    {
      const DualHfvT<F> cx = dual_hfv(0, self);
      if (has[1]) JE = dual_riemann_side(cx, dual_hfv(0, nbr(1)), true);
      if (has[3]) JW = dual_riemann_side(dual_hfv(0, nbr(3)), cx, false);
    }
    {
      const DualHfvT<F> cy = dual_hfv(1, self);
      if (has[2]) JN = dual_riemann_side(cy, dual_hfv(1, nbr(2)), true);
      if (has[4]) JS = dual_riemann_side(dual_hfv(1, nbr(4)), cy, false);
    }
    asm volatile("" ::: "memory");
    // self block without its identity (applied in the storage type by the
    // caller); boundary faces copy the opposite face
    {
      M3T<F> J;
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          const F pE = has[1] ? JE.a[i][j] : JW.a[i][j];
          const F nW = has[3] ? JW.a[i][j] : JE.a[i][j];
          const F pN = has[2] ? JN.a[i][j] : JS.a[i][j];
          const F nS = has[4] ? JS.a[i][j] : JN.a[i][j];
          J.a[i][j] = dtdx * (pE - nW) + dtdy * (pN - nS);
        }
      if (self.thick) {
        const F sint = (F)I.sin_int, tb = (F)mat_of(I, a.mat, a.mtab, c).tan_bed;
        const F alpha = sint * self.sgn[0] * self.k, betta = sint * self.sgn[1] * self.k;
        const F u = self.u, v = self.v;
        const F g2 = self.g2, k0 = stat<F>(a, 5, c), k1 = stat<F>(a, 6, c);
        J.a[1][0] -= dt * (stat<F>(a, 0, c) -
                           alpha * self.h * (F(2) * stat<F>(a, 4, c) + g2 * self.hsy[0]) -
                           alpha * g2 * self.hy);
        if (self.sgn[2] != F(0) && !self.stop[0]) {
          J.a[1][0] += dt * self.sgn[2] * tb * (g2 - u * u * k0);
          J.a[1][1] += dt * self.sgn[2] * tb * (F(2) * u * k0);
        }
        J.a[2][0] -= dt * (stat<F>(a, 1, c) -
                           betta * self.h * (F(2) * stat<F>(a, 3, c) + g2 * self.hsx[0]) -
                           betta * g2 * self.hx);
        if (self.sgn[3] != F(0) && !self.stop[1]) {
          J.a[2][0] += dt * self.sgn[3] * tb * (g2 - v * v * k1);
          J.a[2][2] += dt * self.sgn[3] * tb * (F(2) * v * k1);
        }
      }
      accum_T(adj, J, self.lam);
    }
    const int order[4] = {4, 1, 2, 3};  // S, E, N, W (reference visiting order)
    const F sint = (F)I.sin_int;
#pragma unroll
    for (int o = 0; o < 4; o++) {
      const int s = order[o];
      if (!has[s]) continue;
      const DualInfoT<F> N = nbr(s);
      M3T<F> J = zero3<F>();
      if (s == 1) {
        if (q.i + 2 < I.nx)
          for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) J.a[i][j] = -dtdx * JE.a[i][j];
        nbr_block(J, N, sint, dt, N.hsx[1], F(0));
      } else if (s == 3) {
        if (q.i - 2 >= 0)
          for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) J.a[i][j] = dtdx * JW.a[i][j];
        nbr_block(J, N, sint, dt, N.hsx[2], F(0));
      } else if (s == 2) {
        if (q.j + 2 < I.ny)
          for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) J.a[i][j] = -dtdy * JN.a[i][j];
        nbr_block(J, N, sint, dt, F(0), N.hsy[1]);
      } else {
        if (q.j - 2 >= 0)
          for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) J.a[i][j] = dtdy * JS.a[i][j];
        nbr_block(J, N, sint, dt, F(0), N.hsy[2]);
      }
      accum_T(adj, J, N.lam);
    }
  }
  if (!want_dphi) return 0.;
  asm volatile("" ::: "memory");
  return (double)dphi_term<S, F>(a, I, c, self, nbr(1), nbr(2), nbr(3), nbr(4), has, dt,
                                 self.lam);
}

// Adjoint pass 1 (one workgroup per tile): a tile of void cells (h == 0)
// gets lambda^s = 0 -- every block that multiplies a void cell's state is
// zero (calc_jacobain.C:117-122), as in pass 2 -- and, with lambda^{s+1} == 0
// on the tile, zero bed-friction terms; it is written here.  Every other
// tile is appended to the list pass 2 walks.  With marks, a tile whose U^s
// (and lambda^{s+1}, when the gradient is wanted) is marked +0.0 is void
// without reading the fields, and its lambda^s is not rewritten when that
// is marked +0.0 already.
template <class S, int TX>
__global__ __launch_bounds__(BLOCK) void k32_adjoint_classify(AdjArgsM<S> a) {
  const TileDev t = a.tiles[blockIdx.x];
  const InstDev I = a.inst[t.inst];
  const Cell q = cell_of<TX>(t, I);
  if (a.mk_out && tile_marked_zero<TX>(a.mk_U, t, a.u_bytes) &&
      (!a.dphi_part || tile_marked_zero<TX>(a.mk_lnext, t, a.l_bytes))) {
    if (tile_marked_zero<TX>(a.mk_out, t, a.l_bytes)) return;
    if (q.act) {
      a.lam_h[q.c] = (S)0;
      a.lam_hu[q.c] = (S)0;
      a.lam_hv[q.c] = (S)0;
    }
    tile_mark_store<TX>(a.mk_out, t, true);
    return;
  }
  bool busy = false;
  if (q.act) {
    busy = a.U.h[q.c] != (S)0;
    if (a.dphi_part)
      busy = busy || !(a.lam_next.h[q.c] == (S)0 && a.lam_next.hu[q.c] == (S)0 &&
                       a.lam_next.hv[q.c] == (S)0);
  }
  if (__syncthreads_or(busy)) {
    if (threadIdx.x == 0) a.tile_list[atomicAdd(a.tile_count, 1)] = (int)blockIdx.x;
    return;
  }
  if (q.act) {
    a.lam_h[q.c] = (S)0;
    a.lam_hu[q.c] = (S)0;
    a.lam_hv[q.c] = (S)0;
  }
  if (a.mk_out) tile_mark_store<TX>(a.mk_out, t, true);
}

// Adjoint pass 2, one listed tile.
template <class S, int TX>
__device__ __forceinline__ void adjoint32_tile(const AdjArgsM<S>& a, int tile, DualLds& L,
                                               double* sp) {
  const TileDev t = a.tiles[tile];
  const InstDev I = a.inst[t.inst];
  const Cell q = cell_of<TX>(t, I);
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
  DualInfoT<float> self;
  if (q.act) {
    self = dual_info<S, float>(a, I, q.i, q.j, want_dphi);
    lds_put(L, rself, self);
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
      lds_put(L, ry * RX + rx, dual_info<S, float>(a, I, gi, gj, want_dphi));
  }
  __syncthreads();
  double dphi = 0.;
  bool zero_out = true;  // this cell's lambda^s is +0.0 (or no cell)
  if (q.act) {
    const long long c = q.c;
    const double dtd = a.dt_rec[(long long)t.inst * a.max_steps + a.step];
    // LDS slot of neighbour s (1 E, 2 N, 3 W, 4 S)
    auto rn = [&](int s) {
      return s == 1 ? rself + 1 : (s == 2 ? rself + RX : (s == 3 ? rself - 1 : rself - RX));
    };
    auto thin_at = [&](bool present, int r) {
      const float hn = L.h[r];
      return present && hn > 0.f && hn < THIN;
    };
    // a thin layer in the stencil: the FP64 pass takes the cell
    const bool thin = (self.h > 0.f && self.h < THIN) || thin_at(q.hasE, rn(1)) ||
                      thin_at(q.hasN, rn(2)) || thin_at(q.hasW, rn(3)) || thin_at(q.hasS, rn(4));
    if (thin) {
      // handed to k32_adjoint_thin (FP64), which writes this cell's lambda
      const int slot = atomicAdd(a.thin_count, 1);
      a.thin_list[slot] = make_int2(tile, (int)threadIdx.x);
      zero_out = false;
    } else {
      const S lc[3] = {a.lam_next.h[c], a.lam_next.hu[c], a.lam_next.hv[c]};
      float adj[3];
      dphi = gather<S, float>(a, I, q, self, [&](int s) { return lds_get(L, rn(s)); }, dtd, adj,
                              want_dphi);
      // lambda^s = -(K_self - I)^T lambda_c - sum K_n^T lambda_n = lambda_c - adj;
      // flush |lambda| < 1e-16 (calc_adjoint.C:150-152)
      S out[3];
      for (int j = 0; j < 3; j++) {
        out[j] = self.nz ? lc[j] - (S)adj[j] : (S)0;
        if (fabs((double)out[j]) < 1e-16) out[j] = (S)0;
      }
      a.lam_h[c] = out[0];
      a.lam_hu[c] = out[1];
      a.lam_hv[c] = out[2];
      zero_out = bit_zero(out[0]) && bit_zero(out[1]) && bit_zero(out[2]);
    }
  }
  if (a.mk_out) tile_mark_store<TX>(a.mk_out, t, __syncthreads_and(zero_out));
  if (want_dphi) {
    dphi = wave_sum(dphi);
    if (threadIdx.x % 64 == 0) sp[threadIdx.x / 64] = dphi;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < BLOCK / 64; k++) dphi += sp[k];
      a.dphi_part[tile] += dphi;
    }
  }
}


// Five waves per SIMD: the register cap costs a few spills to scratch, but
// the gather is latency-bound and the extra wave in flight hides more of it
// (about 6% faster than the unconstrained 4-wave allocation, measured).
template <class S, int TX>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5)))
void k32_adjoint(AdjArgsM<S> a) {
  __shared__ DualLds L;
  __shared__ double sp[BLOCK / 64];
  const int n = *a.tile_count;
  for (int i = (int)blockIdx.x; i < n; i += (int)gridDim.x) {
    adjoint32_tile<S, TX>(a, a.tile_list[list_band(i, n)], L, sp);
    __syncthreads();  // LDS reused by the next tile
  }
}

// The cells k32_adjoint listed (a thin layer in their stencil), gathered in
// FP64 from global memory.  A fixed grid strides over the list, so the
// launch does not depend on the count (graph capture).
template <class S, int TX>
__global__ __launch_bounds__(BLOCK) void k32_adjoint_thin(AdjArgsM<S> a) {
  const bool want_dphi = a.dphi_part != nullptr;
  const int n = *a.thin_count;
  for (int e = blockIdx.x * BLOCK + threadIdx.x; e < n; e += gridDim.x * BLOCK) {
    const int2 item = a.thin_list[e];
    const TileDev t = a.tiles[item.x];
    const InstDev I = a.inst[t.inst];
    Cell q;
    q.i = t.x0 + item.y % TX;
    q.j = t.y0 + item.y / TX;
    q.act = true;
    q.c = I.off + (long long)q.j * I.nx + q.i;
    q.hasE = q.i + 1 < I.nx;
    q.hasW = q.i > 0;
    q.hasN = q.j + 1 < I.ny;
    q.hasS = q.j > 0;
    const long long c = q.c;
    const double dtd = a.dt_rec[(long long)t.inst * a.max_steps + a.step];
    const DualInfoT<double> self = dual_info<S, double>(a, I, q.i, q.j, want_dphi);
    double adj[3];
    // neighbour s (1 E, 2 N, 3 W, 4 S); an absent neighbour (domain edge)
    // is never read by the gather -- return the cell's own record instead of
    // touching memory outside the block
    auto nbr = [&](int s) {
      const bool present = s == 1 ? q.hasE : (s == 2 ? q.hasN : (s == 3 ? q.hasW : q.hasS));
      if (!present) return self;
      const int i = q.i + (s == 1 ? 1 : (s == 3 ? -1 : 0));
      const int j = q.j + (s == 2 ? 1 : (s == 4 ? -1 : 0));
      return dual_info<S, double>(a, I, i, j, want_dphi);
    };
    const double dphi = gather<S, double>(a, I, q, self, nbr, dtd, adj, want_dphi);
    const S lc[3] = {a.lam_next.h[c], a.lam_next.hu[c], a.lam_next.hv[c]};
    S out[3];
    for (int j = 0; j < 3; j++) {
      out[j] = self.nz ? (S)((double)lc[j] - adj[j]) : (S)0;
      if (fabs((double)out[j]) < 1e-16) out[j] = (S)0;
    }
    a.lam_h[c] = out[0];
    a.lam_hu[c] = out[1];
    a.lam_hv[c] = out[2];
    if (want_dphi) atomicAdd(&a.dphi_part[item.x], dphi);
  }
}

}  // namespace

#define TG_LAUNCH_TX(shape, KERNEL, S, GRID, ...)                                       \
  do {                                                                                \
    switch (SHAPE_TX[shape]) {                                                        \
      case 64: hipLaunchKernelGGL((KERNEL<S, 64>), GRID, dim3(BLOCK), 0, __VA_ARGS__); break; \
      case 32: hipLaunchKernelGGL((KERNEL<S, 32>), GRID, dim3(BLOCK), 0, __VA_ARGS__); break; \
      default: hipLaunchKernelGGL((KERNEL<S, 16>), GRID, dim3(BLOCK), 0, __VA_ARGS__); break; \
    }                                                                                 \
  } while (0)

template <class S>
void launch_coef_m(const StepArgsM<S>& a, bool reduce, hipStream_t s) {
  TG_LAUNCH_TX(a.shape, k32_coef, S, dim3(a.ntiles), s, a, reduce ? 1 : 0);
}
template <class S>
void launch_update_m(const StepArgsM<S>& a, hipStream_t s) {
  TG_LAUNCH_TX(a.shape, k32_update, S, dim3(a.ntiles), s, a);
}
template <class S>
void launch_adjoint_m(const AdjArgsM<S>& a, hipStream_t s) {
  launch_zero(a.thin_count, sizeof(int), s);
  launch_zero(a.tile_count, sizeof(int), s);
  TG_LAUNCH_TX(a.shape, k32_adjoint_classify, S, dim3(a.ntiles), s, a);
  TG_LAUNCH_TX(a.shape, k32_adjoint, S, dim3(a.ntiles < LIST_GRID ? a.ntiles : LIST_GRID), s,
               a);
  const int g = a.ntiles < THIN_GRID ? a.ntiles : THIN_GRID;
  TG_LAUNCH_TX(a.shape, k32_adjoint_thin, S, dim3(g), s, a);
}
template void launch_coef_m<float>(const StepArgsM<float>&, bool, hipStream_t);
template void launch_coef_m<double>(const StepArgsM<double>&, bool, hipStream_t);
template void launch_update_m<float>(const StepArgsM<float>&, hipStream_t);
template void launch_update_m<double>(const StepArgsM<double>&, hipStream_t);
template void launch_adjoint_m<float>(const AdjArgsM<float>&, hipStream_t);
template void launch_adjoint_m<double>(const AdjArgsM<double>&, hipStream_t);

void launch_stats32(const InstDev* inst, const TileDev* tiles, int ntiles, FieldsT<float> U,
                    double* part_mass, double* part_hmax, double tiny, hipStream_t s) {
  hipLaunchKernelGGL(k32_stats, dim3(ntiles), dim3(BLOCK), 0, s, inst, tiles, U, part_mass,
                     part_hmax, tiny);
}

}  // namespace tg
