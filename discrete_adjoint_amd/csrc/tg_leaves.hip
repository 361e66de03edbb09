// FP64 kernels of the forward step and the per-step discrete adjoint on an
// adaptive mesh given as leaf tables (one thread per leaf, the caller's leaf
// order).  A side of a leaf is a domain boundary, borders one leaf of the
// same generation, one coarser leaf, or two finer leaves; the face rules
// are those of calc_edge_states (element2.C:2744-3072) and
// DualElem::calc_flux (dualmesh.C:1853-2279), written per leaf:
//   * a face between two leaves carries HLL(minus-side leaf, plus-side leaf);
//   * the side of a leaf with two finer neighbours carries half the sum of
//     the two small-face fluxes (the neighbour in slot s first), and its
//     FluxJac entries are (J1 + J2) * .5 for the leaf itself and
//     (J_k + 0) * .5 for each finer neighbour (FluxJac::set, matrix.h:228-234);
//   * a domain-boundary side copies the opposite side (outflow).
// Forward step: k_leaf_coef (slopes, kact, CFL reduction), k_dt,
// k_leaf_update (slopes, face fluxes, update_states).  Reverse step:
// k_leaf_dualk (kact of the dual slopes), k_leaf_jac (the nine Jacobian
// blocks of calc_jacobian, each times lambda^{s+1} of its leaf, stored where
// the leaf that gathers it reads), k_leaf_dphi (bed-friction gradient term,
// when asked for), k_leaf_gather (calc_adjoint's sum, in slot order).
#include "tg_kernels.h"
#include "tg_leafops.h"
#include "tg_math.h"

namespace tg {

namespace {

__device__ __forceinline__ unsigned long long lbits(double v) {
  return (unsigned long long)__double_as_longlong(v);
}
__device__ __forceinline__ double lwave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o, 64);
    v = (w < v) ? w : v;
  }
  return v;
}
__device__ __forceinline__ double lwave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o, 64);
    v = (w > v) ? w : v;
  }
  return v;
}
__device__ __forceinline__ double lwave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ bool lzero3(double a, double b, double c) {
  return (lbits(a) | lbits(b) | lbits(c)) == 0ull;
}
// chunk q and every chunk holding a neighbour of its leaves are +0.0 in mk
__device__ __forceinline__ bool chunk_nbhd_zero(const uint32_t* mk, const LeafDev& L, int q) {
  if (mk[q] != MARK_FULL) return false;
  for (int i = L.cnb_off[q]; i < L.cnb_off[q + 1]; i++)
    if (mk[L.cnb[i]] != MARK_FULL) return false;
  return true;
}
// the writer of chunk q records whether it now holds +0.0 only (every
// thread calls it; all: this thread's leaf is +0.0 or beyond n)
__device__ __forceinline__ void chunk_mark(uint32_t* mk, bool z) {
  const bool all = __syncthreads_and(z);
  if (threadIdx.x == 0) mk[blockIdx.x] = all ? MARK_FULL : 0u;
}

// ---------------------------------------------------------------- forward --

__global__ __launch_bounds__(BLOCK) void k_leaf_coef(StepArgs a, LeafDev L, int reduce) {
  // a chunk of +0.0 leaves (h = 0 < tiny) writes no kact and adds nothing
  if (a.mk_in && a.mk_in[blockIdx.x] == MARK_FULL) return;
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  const InstDev I = a.inst[0];
  double minev = HUGE_VAL, hmax = 0.;
  bool wet = false;
  if (c < L.n) {
    const double h = a.U.h[c];
    wet = h > a.tiny;
    if (wet) {
      const uint32_t meta = L.meta[c];
      const double hu = a.U.hu[c], hv = a.U.hv[c];
      double hx = 0., hux = 0., hy = 0., hvy = 0.;
      if (!touches_boundary(meta)) {
        const DirNb qx = dir_nb(L, c, meta, 0), qy = dir_nb(L, c, meta, 1);
        hx = lslope_fwd(a.U.h, c, qx, a.gamma, a.tiny);
        hux = lslope_fwd(a.U.hu, c, qx, a.gamma, a.tiny);
        hy = lslope_fwd(a.U.h, c, qy, a.gamma, a.tiny);
        hvy = lslope_fwd(a.U.hv, c, qy, a.gamma, a.tiny);
      }
      const double k = getcoef(h, hu, hv, hx, hux, hy, hvy, I.kc1,
                               mat_of(I, a.mat, a.mtab, c).kroot, I.eps, a.tiny);
      if (reduce) {
        double ev = eigen(h, hu, hv, k, a.S.gz[c], a.tiny);
        if (ev > 1000000000.) atomicOr(a.err, 2);
        const double dx = L.cdx[c], dy = L.cdx[L.n + c];
        const double mdx = (dx > dy) ? dy : dx;
        minev = mdx / ev;
        hmax = h;
      }
      a.kact[c] = k;
    } else if (!(h < a.tiny)) {
      // kact is read only where h >= tiny (zero physical flux below tiny,
      // update_states reads it only above): epsilon at tiny, unwritten below
      a.kact[c] = I.eps;
    }
  }
  if (a.active) {
    const int cnt = __syncthreads_count(wet);
    if (threadIdx.x == 0 && cnt) atomicAdd(a.active, (unsigned long long)cnt);
  }
  if (a.short_cells) {
    const int cnt = __syncthreads_count(c < L.n && !wet);
    if (threadIdx.x == 0 && cnt) atomicAdd(a.short_cells, (unsigned long long)cnt);
  }
  if (reduce && __syncthreads_or(wet)) {
    __shared__ double smin[BLOCK / 64], smax[BLOCK / 64];
    minev = lwave_min(minev);
    hmax = lwave_max(hmax);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) { smin[w] = minev; smax[w] = hmax; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < BLOCK / 64; k++) {
        minev = (smin[k] < minev) ? smin[k] : minev;
        hmax = (smax[k] > hmax) ? smax[k] : hmax;
      }
      if (minev < HUGE_VAL) atomicMin(&a.red[0].min_ev, lbits(minev));
      if (hmax > 0.) atomicMax(&a.red[0].hmax, lbits(hmax));
    }
  }
}


// A leaf whose own h and its neighbours' are all below tiny has zero face
// fluxes and no sources: update_states' own arithmetic with zero fluxes.
__device__ __forceinline__ bool leaf_dry(const StepArgs& a, const LeafDev& L, int c,
                                        uint32_t meta) {
  if (!(a.U.h[c] < a.tiny)) return false;
#pragma unroll
  for (int k = 0; k < 8; k++)
    if (slot_present(meta, k) && !(a.U.h[L.nbr[k * L.n + c]] < a.tiny)) return false;
  return true;
}

// One leaf of the update: update_states with its face fluxes.
__device__ __forceinline__ void leaf_update_full(const StepArgs& a, const LeafDev& L, int c,
                                                 uint32_t meta, double out[3]) {
  const InstDev I = a.inst[0];
  const double h = a.U.h[c], hu = a.U.hu[c], hv = a.U.hv[c];
  double hx = 0., hvx = 0., hy = 0., huy = 0.;
  if (!touches_boundary(meta)) {
    const DirNb qx = dir_nb(L, c, meta, 0), qy = dir_nb(L, c, meta, 1);
    hx = lslope_fwd(a.U.h, c, qx, a.gamma, a.tiny);
    hvx = lslope_fwd(a.U.hv, c, qx, a.gamma, a.tiny);
    hy = lslope_fwd(a.U.h, c, qy, a.gamma, a.tiny);
    huy = lslope_fwd(a.U.hu, c, qy, a.gamma, a.tiny);
  }
  const double k = a.kact[c], g2 = a.S.gz[c];
  const Hfv mx = hfv_fwd(0, h, hu, hv, k, g2, a.tiny);
  const Hfv my = hfv_fwd(1, h, hu, hv, k, g2, a.tiny);
  double F[4][3];
#pragma unroll
  for (int s = 0; s < 4; s++)
    if (kind_of(meta, s) != LEAF_BOUNDARY)
      side_flux(L, a.U, a.kact, a.S.gz, a.tiny, c, meta, s, (s & 1) ? mx : my, F[s]);
#pragma unroll
  for (int s = 0; s < 4; s++)  // outflow copies (Element::boundary_flux)
    if (kind_of(meta, s) == LEAF_BOUNDARY)
#pragma unroll
      for (int v = 0; v < 3; v++) F[s][v] = F[(s + 2) & 3][v];
  const double dt = a.dt_rec[a.step];
  const double dtdx = dt / L.cdx[c], dtdy = dt / L.cdx[L.n + c];
  const double prev[3] = {h, hu, hv};
  SrcInfo si;
  update_states(prev, out, F[1], F[2], F[3], F[0], dtdx, dtdy, dt, hx, hvx, hy, huy,
                a.S.kx[c], a.S.ky[c], I.sin_int, mat_of(I, a.mat, a.mtab, c).tan_bed,
                a.S.gx[c], a.S.gy[c], g2,
                a.S.dgx[c], a.S.dgy[c], k, a.ft, a.tiny, &si);
}

// A chunk whose leaves and neighbour leaves are all +0.0 has U^{s+1} =
// U^s - dtdx*(0 - 0) - dtdy*(0 - 0) = +0.0 (finite dt): skipped when its
// output chunk already holds +0.0.
__global__ __launch_bounds__(BLOCK) void k_leaf_update(StepArgs a, LeafDev L) {
  const double dt = a.dt_rec[a.step];
  if (a.mk_in && a.mk_out[blockIdx.x] == MARK_FULL && isfinite(dt) &&
      chunk_nbhd_zero(a.mk_in, L, (int)blockIdx.x))
    return;
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  const uint32_t meta = (c < L.n) ? L.meta[c] : 0u;
  const bool dry = (c < L.n) && leaf_dry(a, L, c, meta);
  if (a.active) {
    const int cnt = __syncthreads_count(c < L.n && !dry);
    if (threadIdx.x == 0 && cnt) atomicAdd(a.active, (unsigned long long)cnt);
  }
  if (a.short_cells) {
    const int cnt = __syncthreads_count(dry);
    if (threadIdx.x == 0 && cnt) atomicAdd(a.short_cells, (unsigned long long)cnt);
  }
  double out[3] = {0., 0., 0.};
  if (dry) {
    const double dtdx = dt / L.cdx[c], dtdy = dt / L.cdx[L.n + c], z = 0.;
    out[0] = a.U.h[c];
    out[1] = a.U.hu[c];
    out[2] = a.U.hv[c];
#pragma unroll
    for (int v = 0; v < 3; v++) out[v] = out[v] - dtdx * (z - z) - dtdy * (z - z);
    if (out[0] < 0.) out[0] = 0;
  } else if (c < L.n) {
    leaf_update_full(a, L, c, meta, out);
  }
  if (c < L.n) {
    if (!isfinite(out[0]) || !isfinite(out[1]) || !isfinite(out[2])) atomicOr(a.err, 1);
    a.h_out[c] = out[0];
    a.hu_out[c] = out[1];
    a.hv_out[c] = out[2];
  }
  if (a.mk_out) chunk_mark(a.mk_out, lzero3(out[0], out[1], out[2]));
}

__global__ __launch_bounds__(BLOCK) void k_leaf_stats(LeafDev L, Fields U, double* part_mass,
                                                      double* part_hmax, double tiny) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  double m = 0., hm = 0.;
  if (c < L.n) {
    const double h = U.h[c];
    m = L.cdx[c] * L.cdx[L.n + c] * h;
    if (h > tiny) hm = h;
  }
  __shared__ double sm[BLOCK / 64], sh[BLOCK / 64];
  m = lwave_sum(m);
  hm = lwave_max(hm);
  if (threadIdx.x % 64 == 0) { sm[threadIdx.x / 64] = m; sh[threadIdx.x / 64] = hm; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < BLOCK / 64; k++) { m += sm[k]; hm = (sh[k] > hm) ? sh[k] : hm; }
    part_mass[blockIdx.x] = m;
    part_hmax[blockIdx.x] = hm;
  }
}

// ---------------------------------------------------------------- adjoint --

// kact of the dual slopes (update_dual_grid, dual_grid_routines.C:68-84) and
// its phi_bed derivative
__device__ __forceinline__ void leaf_dualk_wet(const AdjArgs& a, const LeafDev& L,
                                               const LeafBufs& B, int c, bool want_dk) {
  const InstDev I = a.inst[0];
  const double h = a.U.h[c];
  const uint32_t meta = L.meta[c];
  const double hu = a.U.hu[c], hv = a.U.hv[c];
  double hx = 0., hux = 0., hy = 0., hvy = 0.;
  if (!touches_boundary(meta)) {
    const DirNb qx = dir_nb(L, c, meta, 0), qy = dir_nb(L, c, meta, 1);
    hx = lslope_dual(a.U.h, c, qx, a.gamma, a.tiny, false).d;
    hux = lslope_dual(a.U.hu, c, qx, a.gamma, a.tiny, false).d;
    hy = lslope_dual(a.U.h, c, qy, a.gamma, a.tiny, false).d;
    hvy = lslope_dual(a.U.hv, c, qy, a.gamma, a.tiny, false).d;
  }
  const MatDev M = mat_of(I, a.mat, a.mtab, c);
  B.kd[c] = getcoef(h, hu, hv, hx, hux, hy, hvy, I.kc1, M.kroot, I.eps, a.tiny);
  if (want_dk)
    B.dkd[c] = dgetcoef_dphi(h, hu, hv, hx, hux, hy, hvy, I.kc1, M.kroot, M.dkroot, I.eps,
                             a.tiny);
}

// A chunk of +0.0 leaves whose kd (and dkd, when wanted) already hold the
// dry values is skipped.
__device__ __forceinline__ bool dualk_skip(const AdjArgs& a, const LeafBufs& B, int want_dk) {
  const uint32_t need = want_dk ? 3u : 1u;
  return a.mk_U && a.mk_U[blockIdx.x] == MARK_FULL && (B.kdmk[blockIdx.x] & need) == need;
}
__device__ __forceinline__ void leaf_dualk_chunk(const AdjArgs& a, const LeafDev& L,
                                                 const LeafBufs& B, int want_dk) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  bool dry = true;
  if (c < L.n) {
    const double h = a.U.h[c];
    if (!(h > a.tiny)) {  // getcoef's dry branch (1 * epsilon); no phi_bed dependence
      B.kd[c] = 1. * a.inst[0].eps;
      if (want_dk) B.dkd[c] = 0.;
    } else {
      dry = false;
      leaf_dualk_wet(a, L, B, c, want_dk != 0);
    }
  }
  if (a.mk_U) {
    const bool all = __syncthreads_and(dry);
    if (threadIdx.x == 0) {
      const uint32_t old = B.kdmk[blockIdx.x];
      B.kdmk[blockIdx.x] = all ? (want_dk ? 3u : (1u | (old & 2u))) : 0u;
    }
  }
}
__global__ __launch_bounds__(BLOCK) void k_leaf_dualk(AdjArgs a, LeafDev L, LeafBufs B,
                                                      int want_dk) {
  if (dualk_skip(a, B, want_dk)) return;
  leaf_dualk_chunk(a, L, B, want_dk);
}

__device__ __forceinline__ DualHfv leaf_dual_hfv(const AdjArgs& a, const double* kd, int dir,
                                                 int c) {
  return dual_hfv(dir, a.U.h[c], a.U.hu[c], a.U.hv[c], kd[c], a.S.gz[c], a.tiny);
}

__device__ __forceinline__ M3 half_sum(const M3& x, const M3& y) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.a[i][k] = (x.a[i][k] + y.a[i][k]) * .5;
  return r;
}

// The gather's terms of one block, lam[l] * J[l][k] for every (l, k),
// stored in the record slot of the leaf that adds them.
__device__ __forceinline__ void store_products(double* rec, int n, int slot, int leaf,
                                               const M3& J, const double lam[3]) {
  double* p = rec + ((size_t)slot * n + leaf) * 9;
#pragma unroll
  for (int l = 0; l < 3; l++)
#pragma unroll
    for (int k = 0; k < 3; k++) p[l * 3 + k] = lam[l] * J.a[l][k];
}

// Dual slopes of one leaf (d = (hx, hux, hvx, hy, huy, hvy)) and the h
// slope sensitivities.
__device__ __forceinline__ void dual_slopes(const AdjArgs& a, const LeafDev& L, int c,
                                            uint32_t meta, double d[6], double hs[2][5]) {
#pragma unroll
  for (int k = 0; k < 6; k++) d[k] = 0.;
#pragma unroll
  for (int k = 0; k < 5; k++) hs[0][k] = hs[1][k] = 0.;
  if (touches_boundary(meta)) return;
#pragma unroll
  for (int dir = 0; dir < 2; dir++) {
    const DirNb q = dir_nb(L, c, meta, dir);
    const LSlope sh = lslope_dual(a.U.h, c, q, a.gamma, a.tiny, true);
    d[3 * dir] = sh.d;
#pragma unroll
    for (int k = 0; k < 5; k++) hs[dir][k] = sh.s[k];
    d[3 * dir + 1] = lslope_dual(a.U.hu, c, q, a.gamma, a.tiny, false).d;
    d[3 * dir + 2] = lslope_dual(a.U.hv, c, q, a.gamma, a.tiny, false).d;
  }
}

// Lambda^{s+1}_c . d U^{s+1}_c / d phi_bed at frozen dt (bed-friction
// gradient; no reference counterpart): face fluxes and their kact
// derivatives by the forward face rules.
__device__ __forceinline__ void side_flux_dk(const AdjArgs& a, const LeafDev& L,
                                             const LeafBufs& B, int c, uint32_t meta, int s,
                                             double F[3], double dF[3]) {
  const int dir = (s & 1) ? 0 : 1;
  const bool plus = (s == 1 || s == 2);
  const int cnt = (kind_of(meta, s) == LEAF_FINER) ? 2 : 1;
  double Fi[2][3], dFi[2][3];
  for (int q = 0; q < cnt; q++) {
    const int f = L.nbr[(s + 4 * q) * L.n + c];
    const int l = plus ? c : f, r = plus ? f : c;
    const Hfv hl = hfv_fwd(dir, a.U.h[l], a.U.hu[l], a.U.hv[l], B.kd[l], a.S.gz[l], a.tiny);
    const Hfv hr = hfv_fwd(dir, a.U.h[r], a.U.hu[r], a.U.hv[r], B.kd[r], a.S.gz[r], a.tiny);
    riemann(hl, hr, Fi[q]);
    double dFl[3], dWl[3], dFr[3], dWr[3];
    dhfv_dk(dir, a.U.h[l], B.kd[l], a.S.gz[l], a.tiny, dFl, dWl);
    dhfv_dk(dir, a.U.h[r], B.kd[r], a.S.gz[r], a.tiny, dFr, dWr);
    const double dkl = B.dkd[l], dkr = B.dkd[r];
#pragma unroll
    for (int v = 0; v < 3; v++) {
      dFl[v] *= dkl; dWl[v] *= dkl; dFr[v] *= dkr; dWr[v] *= dkr;
    }
    driemann(hl, hr, dFl, dWl, dFr, dWr, dFi[q]);
  }
#pragma unroll
  for (int v = 0; v < 3; v++) {
    F[v] = (cnt == 2) ? 0.5 * (Fi[0][v] + Fi[1][v]) : Fi[0][v];
    dF[v] = (cnt == 2) ? 0.5 * (dFi[0][v] + dFi[1][v]) : dFi[0][v];
  }
}

// The gradient term of leaf c from the face fluxes F[s] and their phi_bed
// derivatives dF[s] of its four sides s (side_flux_dk; a boundary side's
// entries are replaced by the opposite side's here).
__device__ double leaf_dphi(const AdjArgs& a, const LeafDev& L, const LeafBufs& B,
                            const InstDev& I, int c, uint32_t meta, const double d[6],
                            const SrcInfo& si, double dt, double dtdx, double dtdy,
                            double F[4][3], double dF[4][3]) {
  for (int s = 0; s < 4; s++)
    if (kind_of(meta, s) == LEAF_BOUNDARY)
      for (int v = 0; v < 3; v++) {
        F[s][v] = F[(s + 2) & 3][v];
        dF[s][v] = dF[(s + 2) & 3][v];
      }
  const double h = a.U.h[c], hu = a.U.hu[c], hv = a.U.hv[c];
  double out[3];
#pragma unroll
  for (int v = 0; v < 3; v++)
    out[v] = -dtdx * (dF[1][v] - dF[3][v]) - dtdy * (dF[2][v] - dF[0][v]);
  if (h - dtdx * (F[1][0] - F[3][0]) - dtdy * (F[2][0] - F[0][0]) < 0.) out[0] = 0.;
  if (h > a.tiny) {
    const double tbc = mat_of(I, a.mat, a.mtab, c).tan_bed;
    const double sec2 = 1. + tbc * tbc;
    const double vx = hu / h, vy = hv / h, dk = B.dkd[c];
    const double g2 = a.S.gz[c], dg0 = a.S.dgx[c], dg1 = a.S.dgy[c], k0 = a.S.kx[c],
                 k1 = a.S.ky[c];
    double ds2x = si.sgn[0] * h * dk * (g2 * d[3] + dg1 * h) * I.sin_int;
    double ds3x = si.sgn[2] * umax(g2 * h + vx * hu * k0, 0.0) * sec2;
    double ds2y = si.sgn[1] * h * dk * (g2 * d[0] + dg0 * h) * I.sin_int;
    double ds3y = si.sgn[3] * umax(g2 * h + vy * hv * k1, 0.0) * sec2;
    out[1] += dt * (-ds2x - ds3x);
    out[2] += dt * (-ds2y - ds3y);
  }
  return a.lam_next.h[c] * out[0] + a.lam_next.hu[c] * out[1] + a.lam_next.hv[c] * out[2];
}


// What the blocks of one leaf share (calc_jacobian's per-element values).
struct JacCtx {
  int c, n;
  uint32_t meta;
  double h, dt, dtdx, dtdy, alpha, betta, g2;
  bool wet;
  int stop0, stop1;
  double lam[3];
};

// One neighbour block (set_fluxes_hsens + calc_jacobian_elem for
// effelement = k + 1) from its FluxJac entry E: the own side's entry, and
// the opposite side's -- its copy, so E again -- only when that side is a
// domain boundary.  Stored as gather terms in the neighbour's record.
__device__ __forceinline__ void emit_block(const AdjArgs& a, const LeafDev& L, const LeafBufs& B,
                                           const JacCtx& q, int dir, bool is_plus, bool opp_bnd,
                                           int k, int nb, const M3& E, double dh) {
  // void_neigh_elem (calc_jacobain.C:155-166): a zero block, and a leaf
  // with h == 0 gathers nothing (its lambda is zero), so nothing is stored
  if (a.U.h[nb] == 0.) return;
  const double zero = 0.;
  M3 J;
  {
    const double dh0 = dir ? 0. : dh, dh1 = dir ? dh : 0.;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int r = 0; r < 3; r++) {
        const double e = E.a[i][r];
        const double pp = is_plus ? e : (opp_bnd ? e : zero);
        const double pn = is_plus ? (opp_bnd ? e : zero) : e;
        J.a[i][r] = dir ? (+q.dtdx * (zero - zero) + q.dtdy * (pp - pn))
                        : (+q.dtdx * (pp - pn) + q.dtdy * (zero - zero));
      }
    if (q.wet) {
      if (!q.stop0) J.a[1][0] -= -q.dt * q.h * q.alpha * q.g2 * dh1;
      if (!q.stop1) J.a[2][0] -= -q.dt * q.h * q.betta * q.g2 * dh0;
    }
  }
  store_products(B.jac, q.n, back_of(q.meta, k) + 1, nb, J, q.lam);
}

// One side of one direction: the HLL Jacobians of its face(s) (Case I / II /
// III of DualElem::calc_flux), its neighbour blocks emitted, and its
// FluxJac self entry returned.
__device__ __forceinline__ M3 side_pass(const AdjArgs& a, const LeafDev& L, const LeafBufs& B,
                                        const JacCtx& q, const DualHfv& me, int dir, bool is_plus,
                                        bool opp_bnd, const double hs[5]) {
  const int s = is_plus ? plus_side(dir) : minus_side(dir);
  const int f1 = L.nbr[s * q.n + q.c];
  M3 self1, nb1;
  {
    const DualHfv o1 = leaf_dual_hfv(a, B.kd, dir, f1);
    if (is_plus) dual_riemann_both(me, o1, self1, nb1);
    else dual_riemann_both(o1, me, nb1, self1);
  }
  const double dh1 = hs[is_plus ? 2 : 1], dh2 = hs[is_plus ? 4 : 3];
  if (kind_of(q.meta, s) != LEAF_FINER) {
    emit_block(a, L, B, q, dir, is_plus, opp_bnd, s, f1, nb1, dh1);
    return self1;
  }
  const M3 Z = zero3();
  emit_block(a, L, B, q, dir, is_plus, opp_bnd, s, f1, half_sum(nb1, Z), dh1);
  const int f2 = L.nbr[(s + 4) * q.n + q.c];
  M3 self2, nb2;
  {
    const DualHfv o2 = leaf_dual_hfv(a, B.kd, dir, f2);
    if (is_plus) dual_riemann_both(me, o2, self2, nb2);
    else dual_riemann_both(o2, me, nb2, self2);
  }
  emit_block(a, L, B, q, dir, is_plus, opp_bnd, s + 4, f2, half_sum(nb2, Z), dh2);
  return half_sum(self1, self2);
}

// The nine Jacobian blocks of one leaf (calc_jacobian, calc_jacobain.C:40-153;
// set_fluxes_hsens :187-330; calc_jacobian_elem, compute_jacobian_elem.C:15-121).
// calc_adjoint reads block m of this leaf only from the leaf in slot m - 1,
// where this leaf sits in slot which_neighbor + 1: each block is multiplied
// by this leaf's lambda^{s+1} here and its nine products are stored in that
// reader's record slot, so the gather reads its own records in slot order.
__global__ __launch_bounds__(BLOCK) void k_leaf_jac(AdjArgs a, LeafDev L, LeafBufs B) {
  // void leaves among void neighbours only: no block to store
  if (a.mk_U && chunk_nbhd_zero(a.mk_U, L, (int)blockIdx.x)) return;
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  if (c >= L.n) return;
  const InstDev I = a.inst[0];
  JacCtx q;
  q.c = c;
  q.n = L.n;
  q.meta = L.meta[c];
  const double h = a.U.h[c];
  // A void leaf among void neighbours stores no block (every block it would
  // store lands in the record of a leaf with h == 0).
  if (h == 0.) {
    bool any = false;
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (slot_present(q.meta, k) && a.U.h[L.nbr[k * L.n + c]] != 0.) any = true;
    if (!any) return;
  }
  const double hu = a.U.hu[c], hv = a.U.hv[c];
  q.h = h;
  q.lam[0] = a.lam_next.h[c];
  q.lam[1] = a.lam_next.hu[c];
  q.lam[2] = a.lam_next.hv[c];
  const double kact = B.kd[c];
  const double g0 = a.S.gx[c], g1 = a.S.gy[c], g2 = a.S.gz[c];
  const double dg0 = a.S.dgx[c], dg1 = a.S.dgy[c], cv0 = a.S.kx[c], cv1 = a.S.ky[c];
  // The slopes and source switches enter only the wet self and neighbour
  // terms (h > tiny); a dry leaf's blocks are flux Jacobians alone.
  double d[6] = {0., 0., 0., 0., 0., 0.}, hs[2][5] = {{0., 0., 0., 0., 0.}, {0., 0., 0., 0., 0.}};
  SrcInfo si;
  si.stop[0] = si.stop[1] = 0;
  si.sgn[0] = si.sgn[1] = si.sgn[2] = si.sgn[3] = 0.;
  if (h > a.tiny) {
    dual_slopes(a, L, c, q.meta, d, hs);
    // orgSrcSgn and stop of update_states (they do not depend on the fluxes)
    si = src_switches(h, hu, hv, d[0], d[2], d[3], d[4], cv0, cv1,
                      mat_of(I, a.mat, a.mtab, c).tan_bed, g2, a.ft, a.tiny);
  }
  q.g2 = g2;
  q.dt = a.dt_rec[a.step];
  q.dtdx = q.dt / L.cdx[c];
  q.dtdy = q.dt / L.cdx[q.n + c];
  q.wet = h > a.tiny;
  q.alpha = I.sin_int * si.sgn[0] * kact;
  q.betta = I.sin_int * si.sgn[1] * kact;
  q.stop0 = si.stop[0];
  q.stop1 = si.stop[1];
  double tx[3][3];  // dtdx * (P_x - N_x) of the self block
#pragma unroll
  for (int dir = 0; dir < 2; dir++) {
    const bool bp = kind_of(q.meta, plus_side(dir)) == LEAF_BOUNDARY;
    const bool bm = kind_of(q.meta, minus_side(dir)) == LEAF_BOUNDARY;
    const DualHfv me = leaf_dual_hfv(a, B.kd, dir, c);
    M3 SP = zero3(), SM = zero3();
    if (!bp) SP = side_pass(a, L, B, q, me, dir, true, bm, hs[dir]);
    if (!bm) SM = side_pass(a, L, B, q, me, dir, false, bp, hs[dir]);
    // DualElem::boundary_flux (dualmesh.C:2241-2279): copy the other side
    if (bp) SP = SM;
    if (bm) SM = SP;
    if (dir == 0) {
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) tx[i][k] = q.dtdx * (SP.a[i][k] - SM.a[i][k]);
    } else {
      const double dt = q.dt, alpha = q.alpha, betta = q.betta;
      M3 J;
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) J.a[i][k] = tx[i][k] + q.dtdy * (SP.a[i][k] - SM.a[i][k]);
#pragma unroll
      for (int i = 0; i < 3; i++) J.a[i][i] -= 1.;
      if (q.wet) {
        const double vx = hu / h, vy = hv / h;
        const double vx_sq = vx * vx, vy_sq = vy * vy;
        const double tb = mat_of(I, a.mat, a.mtab, c).tan_bed;
        J.a[1][0] -= dt * (g0 - alpha * h * (2 * dg1 + g2 * hs[1][0]) - alpha * g2 * d[3]);
        if (si.sgn[2] != 0. && !si.stop[0]) {
          J.a[1][0] -= -dt * si.sgn[2] * tb * (g2 - vx_sq * cv0);
          J.a[1][1] -= -dt * si.sgn[2] * tb * (2 * vx * cv0);
          J.a[1][2] -= 0.;
        }
        J.a[2][0] -= dt * (g1 - betta * h * (2 * dg0 + g2 * hs[0][0]) - betta * g2 * d[0]);
        if (si.sgn[3] != 0. && !si.stop[1]) {
          J.a[2][0] -= -dt * si.sgn[3] * tb * (g2 - vy_sq * cv1);
          J.a[2][1] -= 0.;
          J.a[2][2] -= -dt * si.sgn[3] * tb * (2 * vy * cv1);
        }
      }
      if (h != 0.) store_products(B.jac, q.n, 0, c, J, q.lam);
    }
  }
}

// lambda^{s+1}_c . d U^{s+1}_c / d phi_bed at frozen dt, per chunk partial
// sums accumulated over the sweep (bed-friction gradient).  Four lanes per
// leaf (a workgroup takes a quarter of a chunk): lane k evaluates the face
// flux(es) of side k and their phi_bed derivatives -- the sides are
// independent -- and lane 0 gathers the four and adds the source terms, so
// each lane carries a quarter of the face work of the leaf's chain.
constexpr int DPHI_LANES = 4;
__global__ __launch_bounds__(BLOCK) void k_leaf_dphi(AdjArgs a, LeafDev L, LeafBufs B) {
  constexpr int LPW = BLOCK / DPHI_LANES;  // leaves per workgroup
  const int chunk = (int)blockIdx.x / DPHI_LANES;
  // lambda^{s+1} = +0.0 on the whole chunk: zero terms
  if (a.mk_lnext && a.mk_lnext[chunk] == MARK_FULL) return;
  const int c = (int)blockIdx.x * LPW + (int)threadIdx.x / DPHI_LANES;
  const int side = (int)threadIdx.x % DPHI_LANES;
  // lambda^{s+1}_c == 0 contributes a zero term
  const bool act = c < L.n && (a.lam_next.h[c] != 0. || a.lam_next.hu[c] != 0. ||
                               a.lam_next.hv[c] != 0.);
  const uint32_t meta = act ? L.meta[c] : 0u;
  double F1[3] = {0., 0., 0.}, dF1[3] = {0., 0., 0.};
  if (act && kind_of(meta, side) != LEAF_BOUNDARY) side_flux_dk(a, L, B, c, meta, side, F1, dF1);
  // the leaf's four lanes are adjacent in the wave: lane 0 of the quad reads
  // the other three
  double F[4][3], dF[4][3];
  const int q0 = (int)(threadIdx.x % 64) & ~(DPHI_LANES - 1);
#pragma unroll
  for (int k = 0; k < DPHI_LANES; k++)
#pragma unroll
    for (int v = 0; v < 3; v++) {
      F[k][v] = __shfl(F1[v], q0 + k, 64);
      dF[k][v] = __shfl(dF1[v], q0 + k, 64);
    }
  double dphi = 0.;
  if (act && side == 0) {
    const InstDev I = a.inst[0];
    double d[6], hs[2][5];
    dual_slopes(a, L, c, meta, d, hs);
    const SrcInfo si = src_switches(a.U.h[c], a.U.hu[c], a.U.hv[c], d[0], d[2], d[3], d[4],
                                    a.S.kx[c], a.S.ky[c], mat_of(I, a.mat, a.mtab, c).tan_bed,
                                    a.S.gz[c], a.ft, a.tiny);
    const double dt = a.dt_rec[a.step];
    const double dtdx = dt / L.cdx[c], dtdy = dt / L.cdx[L.n + c];
    dphi = leaf_dphi(a, L, B, I, c, meta, d, si, dt, dtdx, dtdy, F, dF);
  }
  __shared__ double sp[BLOCK / 64];
  dphi = lwave_sum(dphi);
  if (threadIdx.x % 64 == 0) sp[threadIdx.x / 64] = dphi;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < BLOCK / 64; k++) dphi += sp[k];
    atomicAdd(&a.dphi_part[chunk], dphi);  // the chunk's DPHI_LANES workgroups
  }
}

// calc_adjoint_elem (calc_adjoint.C:76-157): the terms of slots 0..8 in
// order, lambda_src[l] * J_src[l][k] added in l order for each k
__device__ __forceinline__ void leaf_gather_full(const AdjArgs& a, const LeafDev& L,
                                                 const LeafBufs& B, int c, double out[3]) {
  const int n = L.n;
  const uint32_t meta = L.meta[c];
  double adj[3] = {0., 0., 0.};
#pragma unroll
  for (int m = 0; m < 9; m++) {
    if (m > 0 && !slot_present(meta, m - 1)) continue;
    const double* P = B.jac + ((size_t)m * n + c) * 9;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      adj[k] += P[0 * 3 + k];
      adj[k] += P[1 * 3 + k];
      adj[k] += P[2 * 3 + k];
    }
  }
#pragma unroll
  for (int j = 0; j < 3; j++) {
    double v = 0. - adj[j];
    if (fabs(v) < 1e-16) v = 0.;
    out[j] = v;
  }
}

// calc_adjoint_elem (calc_adjoint.C:76-157) per leaf; a leaf with h == 0
// has every block zero (void_neigh_elem) and lambda = 0, none of its record
// slots written; a chunk of +0.0 leaves whose output already holds +0.0 is
// skipped.
__device__ __forceinline__ bool gather_skip(const AdjArgs& a) {
  return a.mk_U && a.mk_U[blockIdx.x] == MARK_FULL && a.mk_out[blockIdx.x] == MARK_FULL;
}
__device__ __forceinline__ void leaf_gather_chunk(const AdjArgs& a, const LeafDev& L,
                                                  const LeafBufs& B) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  const bool full = c < L.n && a.U.h[c] != 0.;
  if (a.active) {  // leaves on the full path: h != 0
    const int cnt = __syncthreads_count(full);
    if (threadIdx.x == 0 && cnt) atomicAdd(a.active, (unsigned long long)cnt);
  }
  if (a.short_cells) {
    const int cnt = __syncthreads_count(c < L.n && !full);
    if (threadIdx.x == 0 && cnt) atomicAdd(a.short_cells, (unsigned long long)cnt);
  }
  double out[3] = {0., 0., 0.};
  if (full) leaf_gather_full(a, L, B, c, out);
  if (c < L.n) {
    a.lam_h[c] = out[0];
    a.lam_hu[c] = out[1];
    a.lam_hv[c] = out[2];
  }
  if (a.mk_out) chunk_mark(a.mk_out, lzero3(out[0], out[1], out[2]));
}
__global__ __launch_bounds__(BLOCK) void k_leaf_gather(AdjArgs a, LeafDev L, LeafBufs B) {
  if (gather_skip(a)) return;
  leaf_gather_chunk(a, L, B);
}
// The gather of step s (a) and the dual kact of step s - 1 (an: U^{s-1})
// in one launch: the gather reads the step's product records, the dual
// kact writes kd / dkd, which step s's passes have finished reading.
__global__ __launch_bounds__(BLOCK) void k_leaf_gather_dualk(AdjArgs a, AdjArgs an, LeafDev L,
                                                             LeafBufs B, int want_dk) {
  const bool g = !gather_skip(a), d = !dualk_skip(an, B, want_dk);
  if (g) leaf_gather_chunk(a, L, B);
  if (d) leaf_dualk_chunk(an, L, B, want_dk);
}

unsigned leaf_grid(int n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

}  // namespace

void launch_leaf_coef(const StepArgs& a, const LeafDev& L, bool reduce, hipStream_t s) {
  hipLaunchKernelGGL(k_leaf_coef, dim3(leaf_grid(L.n)), dim3(BLOCK), 0, s, a, L, reduce ? 1 : 0);
}

void launch_leaf_update(const StepArgs& a, const LeafDev& L, hipStream_t s) {
  hipLaunchKernelGGL(k_leaf_update, dim3(leaf_grid(L.n)), dim3(BLOCK), 0, s, a, L);
}

void launch_leaf_adjoint(const AdjArgs& a, const AdjArgs* next, bool dualk,
                         const LeafDev& L, const LeafBufs& B, hipStream_t s) {
  const dim3 g(leaf_grid(L.n)), b(BLOCK);
  const int want_dk = a.dphi_part ? 1 : 0;
  if (dualk) hipLaunchKernelGGL(k_leaf_dualk, g, b, 0, s, a, L, B, want_dk);
  hipLaunchKernelGGL(k_leaf_jac, g, b, 0, s, a, L, B);
  if (a.dphi_part)
    hipLaunchKernelGGL(k_leaf_dphi, dim3(leaf_grid(L.n) * DPHI_LANES), b, 0, s, a, L, B);
  if (next) hipLaunchKernelGGL(k_leaf_gather_dualk, g, b, 0, s, a, *next, L, B, want_dk);
  else hipLaunchKernelGGL(k_leaf_gather, g, b, 0, s, a, L, B);
}

void launch_leaf_stats(const LeafDev& L, Fields U, double* part_mass, double* part_hmax,
                       double tiny, hipStream_t s) {
  hipLaunchKernelGGL(k_leaf_stats, dim3(leaf_grid(L.n)), dim3(BLOCK), 0, s, L, U, part_mass,
                     part_hmax, tiny);
}

}  // namespace tg
