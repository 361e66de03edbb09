// Per-leaf operations on the leaf tables (LeafDev) shared by the kernels of
// the adaptive-mesh path and of the error grid: side kinds, the slopes'
// neighbours, forward and dual limited slopes, and the flux of one side.
#pragma once

#include "tg_kernels.h"
#include "tg_math.h"

namespace tg {
namespace {

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

}  // namespace
}  // namespace tg
