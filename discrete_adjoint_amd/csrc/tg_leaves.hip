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
// blocks of calc_jacobian, stored), k_leaf_gather (calc_adjoint's gather
// through which_neighbor).
#include "tg_kernels.h"
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

__device__ __forceinline__ uint32_t kind_of(uint32_t meta, int s) { return (meta >> (2 * s)) & 3u; }
__device__ __forceinline__ int back_of(uint32_t meta, int k) { return (int)((meta >> (8 + 3 * k)) & 7u); }
__device__ __forceinline__ bool slot_present(uint32_t meta, int k) {
  return (k < 4) ? kind_of(meta, k) != LEAF_BOUNDARY : kind_of(meta, k - 4) == LEAF_FINER;
}
// Element::get_slopes' boundary test (element2.C:831-842)
__device__ __forceinline__ bool touches_boundary(uint32_t meta) {
  return kind_of(meta, 0) == LEAF_BOUNDARY || kind_of(meta, 1) == LEAF_BOUNDARY ||
         kind_of(meta, 2) == LEAF_BOUNDARY || kind_of(meta, 3) == LEAF_BOUNDARY;
}

// positive_x_side = 1: x+ side 1, x- side 3, y+ side 2, y- side 0
__device__ __forceinline__ int plus_side(int dir) { return dir ? 2 : 1; }
__device__ __forceinline__ int minus_side(int dir) { return dir ? 0 : 3; }

// The neighbours one direction's slopes read: first and (on a side with two
// finer leaves) second neighbour of each side, and the centre distances to
// the first ones (element2.C:866-895).
struct DirNb {
  int p, m, p2, m2;
  double dxp, dxm;
};
__device__ __forceinline__ DirNb dir_nb(const LeafDev& L, int c, uint32_t meta, int dir) {
  const int sp = plus_side(dir), sm = minus_side(dir), n = L.n;
  DirNb q;
  q.p = L.nbr[sp * n + c];
  q.m = L.nbr[sm * n + c];
  q.p2 = (kind_of(meta, sp) == LEAF_FINER) ? L.nbr[(sp + 4) * n + c] : -1;
  q.m2 = (kind_of(meta, sm) == LEAF_FINER) ? L.nbr[(sm + 4) * n + c] : -1;
  const double* X = L.xc + (size_t)dir * n;
  const double xc = X[c];
  q.dxp = X[q.p] - xc;
  q.dxm = xc - X[q.m];
  return q;
}

// Element::get_slopes, element2.C:886-897 (divides by the distances)
__device__ __forceinline__ double lslope_fwd(const double* u, int c, const DirNb& q,
                                             double gamma, double tiny) {
  const double uc = u[c];
  double dp = (u[q.p] - uc) / q.dxp;
  if (q.p2 >= 0) dp = .5 * (dp + (u[q.p2] - uc) / q.dxp);
  double dm = (uc - u[q.m]) / q.dxm;
  if (q.m2 >= 0) dm = .5 * (dm + (uc - u[q.m2]) / q.dxm);
  double dc = (dp * q.dxm + dm * q.dxp) / (q.dxm + q.dxp);
  return .5 * (c_sgn(dp, tiny) + c_sgn(dm, tiny)) *
         dmin3(gamma * dabs(dp), gamma * dabs(dm), dabs(dc));
}

// DualElem::get_slopes_prev, dualmesh.C:1457-1532 (multiplies by
// reciprocals; hslope_sens {self, minus1, plus1, minus2, plus2} for h)
struct LSlope {
  double d;
  double s[5];
};
__device__ __forceinline__ LSlope lslope_dual(const double* u, int c, const DirNb& q,
                                              double gamma, double tiny, bool sens) {
  LSlope r;
  const double inv_dxp = 1 / q.dxp, inv_dxm = 1 / q.dxm, inv_s = 1 / (q.dxm + q.dxp);
  const double uc = u[c];
  double dp = (u[q.p] - uc) * inv_dxp;
  if (q.p2 >= 0) dp = .5 * (dp + (u[q.p2] - uc) * inv_dxp);
  double dm = (uc - u[q.m]) * inv_dxm;
  if (q.m2 >= 0) dm = .5 * (dm + (uc - u[q.m2]) * inv_dxm);
  double dc = (dp * q.dxm + dm * q.dxp) * inv_s;
  double mn = dmin3(gamma * dabs(dp), gamma * dabs(dm), dabs(dc));
  double sg = .5 * (c_sgn(dp, tiny) + c_sgn(dm, tiny));
  r.d = sg * mn;
#pragma unroll
  for (int k = 0; k < 5; k++) r.s[k] = 0.;
  if (sens) {
    if (sg == 0) {
    } else if (mn == gamma * dabs(dp)) {
      r.s[0] = -inv_dxp;
      r.s[2] = inv_dxp;
      if (q.p2 >= 0) r.s[2] = r.s[4] = .5 * inv_dxp;
    } else if (mn == gamma * dabs(dm)) {
      r.s[0] = inv_dxm;
      r.s[1] = -inv_dxm;
      if (q.m2 >= 0) r.s[1] = r.s[3] = -.5 * inv_dxm;
    } else {
      r.s[0] = inv_dxm - inv_dxp;
      r.s[1] = -inv_dxm;
      r.s[2] = inv_dxp;
      if (q.p2 >= 0) r.s[2] = r.s[4] = .5 * inv_dxp;
      if (q.m2 >= 0) r.s[1] = r.s[3] = -.5 * inv_dxm;
      r.s[1] *= q.dxp * inv_s;
      r.s[2] *= q.dxm * inv_s;
      r.s[3] *= q.dxp * inv_s;
      r.s[4] *= q.dxm * inv_s;
    }
  }
  return r;
}

__device__ __forceinline__ Hfv leaf_hfv(const Fields& U, const double* k, const double* gz,
                                        int dir, int c, double tiny) {
  return hfv_fwd(dir, U.h[c], U.hu[c], U.hv[c], k[c], gz[c], tiny);
}

// Flux of side s of leaf c (not a boundary side): HLL with the minus-side
// leaf on the left; half the sum of both small faces on a side with two
// finer neighbours.
__device__ __forceinline__ void side_flux(const LeafDev& L, const Fields& U, const double* k,
                                          const double* gz, double tiny, int c, uint32_t meta,
                                          int s, const Hfv& me, double F[3]) {
  const int dir = (s & 1) ? 0 : 1;
  const bool plus = (s == 1 || s == 2);
  const Hfv o = leaf_hfv(U, k, gz, dir, L.nbr[s * L.n + c], tiny);
  if (plus) riemann(me, o, F);
  else riemann(o, me, F);
  if (kind_of(meta, s) == LEAF_FINER) {
    const Hfv o2 = leaf_hfv(U, k, gz, dir, L.nbr[(s + 4) * L.n + c], tiny);
    double F2[3];
    if (plus) riemann(me, o2, F2);
    else riemann(o2, me, F2);
#pragma unroll
    for (int v = 0; v < 3; v++) F[v] = 0.5 * (F[v] + F2[v]);
  }
}

// ---------------------------------------------------------------- forward --

__global__ __launch_bounds__(BLOCK) void k_leaf_coef(StepArgs a, LeafDev L, int reduce) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  const InstDev I = a.inst[0];
  double minev = HUGE_VAL, hmax = 0.;
  if (c < L.n) {
    const uint32_t meta = L.meta[c];
    const double h = a.U.h[c], hu = a.U.hu[c], hv = a.U.hv[c];
    double hx = 0., hux = 0., hy = 0., hvy = 0.;
    if (!touches_boundary(meta)) {
      const DirNb qx = dir_nb(L, c, meta, 0), qy = dir_nb(L, c, meta, 1);
      hx = lslope_fwd(a.U.h, c, qx, a.gamma, a.tiny);
      hux = lslope_fwd(a.U.hu, c, qx, a.gamma, a.tiny);
      hy = lslope_fwd(a.U.h, c, qy, a.gamma, a.tiny);
      hvy = lslope_fwd(a.U.hv, c, qy, a.gamma, a.tiny);
    }
    double k;
    if (h > a.tiny) {
      k = getcoef(h, hu, hv, hx, hux, hy, hvy, I.kc1, I.kroot, I.eps, a.tiny);
      if (reduce) {
        double ev = eigen(h, hu, hv, k, a.S.gz[c], a.tiny);
        if (ev > 1000000000.) atomicOr(a.err, 2);
        const double dx = L.cdx[c], dy = L.cdx[L.n + c];
        const double mdx = (dx > dy) ? dy : dx;
        minev = mdx / ev;
        hmax = h;
      }
    } else {
      k = I.eps;
    }
    a.kact[c] = k;
  }
  if (reduce) {
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

__global__ __launch_bounds__(BLOCK) void k_leaf_update(StepArgs a, LeafDev L) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  if (c >= L.n) return;
  const InstDev I = a.inst[0];
  const uint32_t meta = L.meta[c];
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
  double out[3];
  SrcInfo si;
  update_states(prev, out, F[1], F[2], F[3], F[0], dtdx, dtdy, dt, hx, hvx, hy, huy,
                a.S.kx[c], a.S.ky[c], I.sin_int, I.tan_bed, a.S.gx[c], a.S.gy[c], g2,
                a.S.dgx[c], a.S.dgy[c], k, a.ft, a.tiny, &si);
  if (!isfinite(out[0]) || !isfinite(out[1]) || !isfinite(out[2])) atomicOr(a.err, 1);
  a.h_out[c] = out[0];
  a.hu_out[c] = out[1];
  a.hv_out[c] = out[2];
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
__global__ __launch_bounds__(BLOCK) void k_leaf_dualk(AdjArgs a, LeafDev L, LeafBufs B,
                                                      int want_dk) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  if (c >= L.n) return;
  const InstDev I = a.inst[0];
  const uint32_t meta = L.meta[c];
  const double h = a.U.h[c], hu = a.U.hu[c], hv = a.U.hv[c];
  double hx = 0., hux = 0., hy = 0., hvy = 0.;
  if (!touches_boundary(meta)) {
    const DirNb qx = dir_nb(L, c, meta, 0), qy = dir_nb(L, c, meta, 1);
    hx = lslope_dual(a.U.h, c, qx, a.gamma, a.tiny, false).d;
    hux = lslope_dual(a.U.hu, c, qx, a.gamma, a.tiny, false).d;
    hy = lslope_dual(a.U.h, c, qy, a.gamma, a.tiny, false).d;
    hvy = lslope_dual(a.U.hv, c, qy, a.gamma, a.tiny, false).d;
  }
  B.kd[c] = getcoef(h, hu, hv, hx, hux, hy, hvy, I.kc1, I.kroot, I.eps, a.tiny);
  if (want_dk)
    B.dkd[c] = dgetcoef_dphi(h, hu, hv, hx, hux, hy, hvy, I.kc1, I.kroot, I.dkroot, I.eps,
                             a.tiny);
}

__device__ __forceinline__ DualHfv leaf_dual_hfv(const AdjArgs& a, const double* kd, int dir,
                                                 int c) {
  return dual_hfv(dir, a.U.h[c], a.U.hu[c], a.U.hv[c], kd[c], a.S.gz[c], a.tiny);
}

// FluxJac entries of one direction (side x or y): [minus/plus][self/n1/n2]
struct FluxJacDir {
  M3 j[2][3];
};

__device__ __forceinline__ M3 half_sum(const M3& x, const M3& y) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.a[i][k] = (x.a[i][k] + y.a[i][k]) * .5;
  return r;
}

__device__ __forceinline__ FluxJacDir flux_jac_dir(const AdjArgs& a, const LeafDev& L,
                                                   const double* kd, int c, uint32_t meta,
                                                   int dir) {
  FluxJacDir F;
  const DualHfv me = leaf_dual_hfv(a, kd, dir, c);
  const M3 Z = zero3();
#pragma unroll
  for (int pm = 0; pm < 2; pm++) {
    const int s = pm ? plus_side(dir) : minus_side(dir);
    const uint32_t kd_s = kind_of(meta, s);
    if (kd_s == LEAF_BOUNDARY) continue;
    const bool plus = pm == 1;
    const DualHfv o1 = leaf_dual_hfv(a, kd, dir, L.nbr[s * L.n + c]);
    // d HLL / d U of this leaf and of the neighbour (left = minus side)
    const M3 self1 = plus ? dual_riemann_side(me, o1, true) : dual_riemann_side(o1, me, false);
    const M3 nb1 = plus ? dual_riemann_side(me, o1, false) : dual_riemann_side(o1, me, true);
    if (kd_s == LEAF_FINER) {
      const DualHfv o2 = leaf_dual_hfv(a, kd, dir, L.nbr[(s + 4) * L.n + c]);
      const M3 self2 =
          plus ? dual_riemann_side(me, o2, true) : dual_riemann_side(o2, me, false);
      const M3 nb2 = plus ? dual_riemann_side(me, o2, false) : dual_riemann_side(o2, me, true);
      F.j[pm][0] = half_sum(self1, self2);
      F.j[pm][1] = half_sum(nb1, Z);
      F.j[pm][2] = half_sum(nb2, Z);
    } else {
      F.j[pm][0] = self1;
      F.j[pm][1] = nb1;
      F.j[pm][2] = Z;
    }
  }
  // DualElem::boundary_flux (dualmesh.C:2241-2279): copy the other side
  if (kind_of(meta, plus_side(dir)) == LEAF_BOUNDARY)
#pragma unroll
    for (int i = 0; i < 3; i++) F.j[1][i] = F.j[0][i];
  if (kind_of(meta, minus_side(dir)) == LEAF_BOUNDARY)
#pragma unroll
    for (int i = 0; i < 3; i++) F.j[0][i] = F.j[1][i];
  return F;
}

__device__ __forceinline__ void store_block(double* jac, int n, int m, int c, const M3& J) {
  double* p = jac + ((size_t)m * n + c) * 9;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) p[i * 3 + k] = J.a[i][k];
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

__device__ double leaf_dphi(const AdjArgs& a, const LeafDev& L, const LeafBufs& B,
                            const InstDev& I, int c, uint32_t meta, const double d[6],
                            const SrcInfo& si, double dt, double dtdx, double dtdy) {
  double F[4][3], dF[4][3];
  for (int s = 0; s < 4; s++)
    if (kind_of(meta, s) != LEAF_BOUNDARY) side_flux_dk(a, L, B, c, meta, s, F[s], dF[s]);
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
    const double sec2 = 1. + I.tan_bed * I.tan_bed;
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

// The nine Jacobian blocks of one leaf (calc_jacobian, calc_jacobain.C:40-153;
// set_fluxes_hsens :187-330; calc_jacobian_elem, compute_jacobian_elem.C:15-121),
// stored for the gather; blocks of absent slots are not written (the gather
// does not read them).
__global__ __launch_bounds__(BLOCK) void k_leaf_jac(AdjArgs a, LeafDev L, LeafBufs B) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  const bool act = c < L.n;
  double dphi = 0.;
  if (act) {
    const InstDev I = a.inst[0];
    const uint32_t meta = L.meta[c];
    const int n = L.n;
    const double h = a.U.h[c], hu = a.U.hu[c], hv = a.U.hv[c];
    double d[6] = {0., 0., 0., 0., 0., 0.};
    double hs[2][5] = {{0., 0., 0., 0., 0.}, {0., 0., 0., 0., 0.}};
    if (!touches_boundary(meta)) {
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
    const double kact = B.kd[c];
    const double g0 = a.S.gx[c], g1 = a.S.gy[c], g2 = a.S.gz[c];
    const double dg0 = a.S.dgx[c], dg1 = a.S.dgy[c], cv0 = a.S.kx[c], cv1 = a.S.ky[c];
    // orgSrcSgn and stop of update_states (they do not depend on the fluxes)
    const SrcInfo si = src_switches(h, hu, hv, d[0], d[2], d[3], d[4], cv0, cv1, I.tan_bed, g2,
                                    a.ft, a.tiny);
    const double dt = a.dt_rec[a.step];
    const double dtdx = dt / L.cdx[c], dtdy = dt / L.cdx[n + c];
    const bool wet = h > a.tiny;
    const double alpha = I.sin_int * si.sgn[0] * kact;
    const double betta = I.sin_int * si.sgn[1] * kact;
    const double zero = 0.;
    double tx[3][3];  // dtdx * (P_x - N_x) of the self block
#pragma unroll
    for (int dir = 0; dir < 2; dir++) {
      const FluxJacDir FJ = flux_jac_dir(a, L, B.kd, c, meta, dir);
      if (dir == 0) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int k = 0; k < 3; k++) tx[i][k] = dtdx * (FJ.j[1][0].a[i][k] - FJ.j[0][0].a[i][k]);
      } else {
        M3 J;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int k = 0; k < 3; k++)
            J.a[i][k] = tx[i][k] + dtdy * (FJ.j[1][0].a[i][k] - FJ.j[0][0].a[i][k]);
#pragma unroll
        for (int i = 0; i < 3; i++) J.a[i][i] -= 1.;
        if (wet) {
          const double vx = hu / h, vy = hv / h;
          const double vx_sq = vx * vx, vy_sq = vy * vy, tb = I.tan_bed;
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
        store_block(B.jac, n, 0, c, (h == 0.) ? zero3() : J);
      }
      // neighbour slots of this direction: side sp (slot sp+1, sp+5) and sm
      const int sp = plus_side(dir), sm = minus_side(dir);
#pragma unroll
      for (int side2 = 0; side2 < 2; side2++) {
        const int sd = side2 ? sm : sp;
        const bool is_plus = side2 == 0;
        const bool opp_bnd = kind_of(meta, is_plus ? sm : sp) == LEAF_BOUNDARY;
#pragma unroll
        for (int idx = 1; idx <= 2; idx++) {
          const int k = sd + 4 * (idx - 1);  // table slot (block m = k + 1)
          if (!slot_present(meta, k)) continue;
          const int nb = L.nbr[k * n + c];
          M3 J;
          if (a.U.h[nb] == 0.) {  // void_neigh_elem (calc_jacobain.C:155-166)
            J = zero3();
          } else {
            // set_fluxes_hsens: the own side's entry; the opposite side's
            // entry only when that side is a boundary (its copy)
            M3 pm_plus, pm_minus;
            if (is_plus) {
              pm_plus = FJ.j[1][idx];
              pm_minus = opp_bnd ? FJ.j[0][idx] : zero3();
            } else {
              pm_minus = FJ.j[0][idx];
              pm_plus = opp_bnd ? FJ.j[1][idx] : zero3();
            }
            const double dh = hs[dir][is_plus ? (idx == 1 ? 2 : 4) : (idx == 1 ? 1 : 3)];
            const double dh0 = dir ? 0. : dh, dh1 = dir ? dh : 0.;
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
              for (int q = 0; q < 3; q++)
                J.a[i][q] = dir ? (+dtdx * (zero - zero) +
                                   dtdy * (pm_plus.a[i][q] - pm_minus.a[i][q]))
                                : (+dtdx * (pm_plus.a[i][q] - pm_minus.a[i][q]) +
                                   dtdy * (zero - zero));
            if (wet) {
              if (!si.stop[0]) J.a[1][0] -= -dt * h * alpha * g2 * dh1;
              if (!si.stop[1]) J.a[2][0] -= -dt * h * betta * g2 * dh0;
            }
          }
          store_block(B.jac, n, k + 1, c, J);
        }
      }
    }
    if (a.dphi_part) dphi = leaf_dphi(a, L, B, I, c, meta, d, si, dt, dtdx, dtdy);
  }
  if (a.dphi_part) {
    __shared__ double sp[BLOCK / 64];
    dphi = lwave_sum(dphi);
    if (threadIdx.x % 64 == 0) sp[threadIdx.x / 64] = dphi;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < BLOCK / 64; k++) dphi += sp[k];
      a.dphi_part[blockIdx.x] += dphi;
    }
  }
}

// calc_adjoint_elem (calc_adjoint.C:76-157): slots 0..8 in order, each
// neighbour's block that points back at this leaf (which_neighbor + 1)
__global__ __launch_bounds__(BLOCK) void k_leaf_gather(AdjArgs a, LeafDev L, LeafBufs B) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  if (c >= L.n) return;
  const int n = L.n;
  const uint32_t meta = L.meta[c];
  double adj[3] = {0., 0., 0.};
#pragma unroll 1
  for (int m = 0; m < 9; m++) {
    int src = c, blk = 0;
    if (m > 0) {
      if (!slot_present(meta, m - 1)) continue;
      src = L.nbr[(m - 1) * n + c];
      blk = back_of(meta, m - 1) + 1;
    }
    const double* J = B.jac + ((size_t)blk * n + src) * 9;
    const double l0 = a.lam_next.h[src], l1 = a.lam_next.hu[src], l2 = a.lam_next.hv[src];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      adj[k] += l0 * J[0 * 3 + k];
      adj[k] += l1 * J[1 * 3 + k];
      adj[k] += l2 * J[2 * 3 + k];
    }
  }
  double out[3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    double v = 0. - adj[j];
    if (fabs(v) < 1e-16) v = 0.;
    out[j] = v;
  }
  a.lam_h[c] = out[0];
  a.lam_hu[c] = out[1];
  a.lam_hv[c] = out[2];
}

unsigned leaf_grid(int n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

}  // namespace

void launch_leaf_coef(const StepArgs& a, const LeafDev& L, bool reduce, hipStream_t s) {
  hipLaunchKernelGGL(k_leaf_coef, dim3(leaf_grid(L.n)), dim3(BLOCK), 0, s, a, L, reduce ? 1 : 0);
}

void launch_leaf_update(const StepArgs& a, const LeafDev& L, hipStream_t s) {
  hipLaunchKernelGGL(k_leaf_update, dim3(leaf_grid(L.n)), dim3(BLOCK), 0, s, a, L);
}

void launch_leaf_adjoint(const AdjArgs& a, const LeafDev& L, const LeafBufs& B, hipStream_t s) {
  const dim3 g(leaf_grid(L.n)), b(BLOCK);
  hipLaunchKernelGGL(k_leaf_dualk, g, b, 0, s, a, L, B, a.dphi_part ? 1 : 0);
  hipLaunchKernelGGL(k_leaf_jac, g, b, 0, s, a, L, B);
  hipLaunchKernelGGL(k_leaf_gather, g, b, 0, s, a, L, B);
}

void launch_leaf_stats(const LeafDev& L, Fields U, double* part_mass, double* part_hmax,
                       double tiny, hipStream_t s) {
  hipLaunchKernelGGL(k_leaf_stats, dim3(leaf_grid(L.n)), dim3(BLOCK), 0, s, L, U, part_mass,
                     part_hmax, tiny);
}

}  // namespace tg
