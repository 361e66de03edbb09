// Mixed-precision per-cell arithmetic.  Same algorithm as tg_math.h (which
// restates the reference operand-for-operand in FP64), re-associated for
// fewer operations: one reciprocal of h per cell instead of repeated
// divisions, hardware reciprocal / square root.
//
// Two arithmetic types appear:
//  - W, the "switch" type: everything that decides a branch of the
//    reference's non-smooth scheme (limited slopes and their c_sgn dead
//    zone, the slope-limiter branch, 1/h and velocities, the kact sign of the
//    velocity divergence, tiny_sgn of the velocity gradient, the Coulomb
//    clamp).  These arguments are differences of nearly equal terms; in FP32
//    the cancellation flips the switches and the flips, not rounding,
//    dominate the error.  W = double in TG_MIXED, float in TG_FP32.
//  - float for the bulk: fluxes, HLL, sources, Jacobian blocks.
#pragma once
#include <hip/hip_runtime.h>

namespace tg {
namespace m32 {

constexpr float THIN = 1e-3f;

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// FP64 counterparts (IEEE division / square root) for the thin-layer path
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }

template <class W>
__device__ __forceinline__ W c_sgn(W z, W tiny) {
  return (z > tiny) ? W(1) : ((z < -tiny) ? W(-1) : W(0));
}
template <class W>
__device__ __forceinline__ W f_sgn(W z, W tiny) {
  return (fabs(z) < tiny) ? W(0) : ((z >= tiny) ? W(1) : W(-1));
}
template <class W>
__device__ __forceinline__ W tiny_sgn(W n, W t) {
  return (fabs(n) < t) ? W(0) : ((n > t) ? W(1) : W(-1));
}

// 1/h: FP64 division for W = double; FP32 reciprocal, or FP64 for thin
// layers (h < 1e-3, where FP32 velocity recovery loses the most), for float
__device__ __forceinline__ double inv_h(double h) { return (h > 0.) ? 1.0 / h : 0.; }
__device__ __forceinline__ float inv_h(float h) {
  if (!(h > 0.f)) return 0.f;
  return (h < THIN) ? (float)(1.0 / (double)h) : rcp(h);
}
__device__ __forceinline__ double vel(double hu, double h, double ih) { return hu * ih; }
__device__ __forceinline__ float vel(float hu, float h, float ih) {
  return (h < THIN) ? (float)((double)hu / (double)h) : hu * ih;
}

// limited slope on a uniform spacing (dxp == dxm: dc = (dp + dm) / 2)
template <class W>
__device__ __forceinline__ W slope(W up, W uc, W um, W inv_dx, W gamma, W tiny) {
  W dp = (up - uc) * inv_dx, dm = (uc - um) * inv_dx, dc = W(0.5) * (dp + dm);
  W mn = fmin(fmin(gamma * fabs(dp), gamma * fabs(dm)), fabs(dc));
  return W(0.5) * (c_sgn(dp, tiny) + c_sgn(dm, tiny)) * mn;
}

template <class W>
struct Slope {
  W d, s0, s1, s2;  // value and hslope_sens {self, minus, plus}
};
template <class W>
__device__ __forceinline__ Slope<W> slope_sens(W up, W uc, W um, W inv_dx, W gamma, W tiny) {
  Slope<W> r;
  W dp = (up - uc) * inv_dx, dm = (uc - um) * inv_dx, dc = W(0.5) * (dp + dm);
  W ap = gamma * fabs(dp), am = gamma * fabs(dm), ac = fabs(dc);
  W mn = fmin(fmin(ap, am), ac);
  W sg = W(0.5) * (c_sgn(dp, tiny) + c_sgn(dm, tiny));
  r.d = sg * mn;
  r.s0 = r.s1 = r.s2 = W(0);
  if (sg != W(0)) {
    if (mn == ap) { r.s0 = -inv_dx; r.s2 = inv_dx; }
    else if (mn == am) { r.s0 = inv_dx; r.s1 = -inv_dx; }
    else { r.s1 = W(-0.5) * inv_dx; r.s2 = W(0.5) * inv_dx; }
  }
  return r;
}

// getcoef: kact (times epsilon) in arithmetic type F; sign decision in W
template <class W, class F = float>
__device__ __forceinline__ F kact(W h, W u, W v, W ih, W hx, W hux, W hy, W hvy, F kc1, F kroot,
                                  F eps, W tiny) {
  if (!(h > tiny)) return eps;
  W velg = (hux - u * hx + hvy - v * hy) * ih;
  F k = kc1 * (F(1) - (F)f_sgn(velg, tiny) * kroot) - F(1);
  if (fabs(u) < tiny && fabs(v) < tiny) k = F(1);
  return fmax(k, F(0)) * eps;
}

template <class W, class F = float>
__device__ __forceinline__ F dkact_dphi(W h, W u, W v, W ih, W hx, W hux, W hy, W hvy, F kc1,
                                        F kroot, F dkroot, F eps, W tiny) {
  if (!(h > tiny)) return F(0);
  W velg = (hux - u * hx + hvy - v * hy) * ih;
  F s = (F)f_sgn(velg, tiny);
  F k = kc1 * (F(1) - s * kroot) - F(1);
  if (fabs(u) < tiny && fabs(v) < tiny) return F(0);
  if (k < F(0)) return F(0);
  return -kc1 * s * dkroot * eps;
}

// Physical flux of one side, its state and wave speeds, in arithmetic type F
// (float for the bulk, double for the thin-layer path of the adjoint).
template <class F>
struct HfvT {
  F u[3], f[3], w[3];
};
using Hfv = HfvT<float>;

// physical flux of one side in direction dir (0 x, 1 y); un = normal velocity
template <class F>
__device__ __forceinline__ HfvT<F> hfv(int dir, F h, F hu, F hv, F un, F k, F g3, bool wet) {
  HfvT<F> r;
  if (!wet) {
#pragma unroll
    for (int i = 0; i < 3; i++) r.u[i] = r.f[i] = r.w[i] = F(0);
    return r;
  }
  F a2 = k * h * g3, a = fsqrt(a2);
  r.u[0] = h; r.u[1] = hu; r.u[2] = hv;
  r.f[0] = dir ? hv : hu;
  r.f[1] = hu * un + (dir ? F(0) : F(0.5) * a2 * h);
  r.f[2] = hv * un + (dir ? F(0.5) * a2 * h : F(0));
  r.w[0] = un - a; r.w[1] = un; r.w[2] = un + a;
  return r;
}

template <class F>
struct SpeedsT {
  F sl, sr;
  int mode;  // 0 zero flux, 1 left upwind, 2 right upwind, 3 HLL
};
template <class F>
__device__ __forceinline__ SpeedsT<F> speeds(const HfvT<F>& l, const HfvT<F>& r) {
  SpeedsT<F> s;
  if (l.u[0] == F(0) && r.u[0] == F(0)) { s.sl = s.sr = F(0); s.mode = 0; return s; }
  if (l.u[0] == F(0)) {
    s.sl = fmin(F(0), F(2) * r.w[0] - r.w[1]);
    s.sr = fmax(F(0), F(2) * r.w[2] - r.w[1]);
  } else if (r.u[0] == F(0)) {
    s.sl = fmin(F(0), F(2) * l.w[0] - l.w[1]);
    s.sr = fmax(F(0), F(2) * l.w[2] - l.w[1]);
  } else {
    s.sl = fmin(F(0), fmin(l.w[0], r.w[0]));
    s.sr = fmax(F(0), fmax(l.w[2], r.w[2]));
  }
  s.mode = (s.sl >= F(0)) ? 1 : ((s.sr <= F(0)) ? 2 : 3);
  return s;
}

template <class F>
__device__ __forceinline__ void riemann(const HfvT<F>& l, const HfvT<F>& r, F flux[3]) {
  SpeedsT<F> s = speeds(l, r);
  if (s.mode == 0) { flux[0] = flux[1] = flux[2] = F(0); return; }
  if (s.mode == 1) { for (int v = 0; v < 3; v++) flux[v] = l.f[v]; return; }
  if (s.mode == 2) { for (int v = 0; v < 3; v++) flux[v] = r.f[v]; return; }
  F inv = rcp(s.sr - s.sl);
#pragma unroll
  for (int v = 0; v < 3; v++)
    flux[v] = (s.sr * l.f[v] - s.sl * r.f[v] + s.sl * s.sr * (r.u[v] - l.u[v])) * inv;
}

}  // namespace m32
}  // namespace tg
