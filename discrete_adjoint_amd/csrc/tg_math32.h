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

// getcoef: kact (times epsilon); sign decision in W
template <class W>
__device__ __forceinline__ float kact(W h, W u, W v, W ih, W hx, W hux, W hy, W hvy, float kc1,
                                      float kroot, float eps, W tiny) {
  if (!(h > tiny)) return eps;
  W velg = (hux - u * hx + hvy - v * hy) * ih;
  float k = kc1 * (1.f - (float)f_sgn(velg, tiny) * kroot) - 1.f;
  if (fabs(u) < tiny && fabs(v) < tiny) k = 1.f;
  return fmaxf(k, 0.f) * eps;
}

template <class W>
__device__ __forceinline__ float dkact_dphi(W h, W u, W v, W ih, W hx, W hux, W hy, W hvy,
                                            float kc1, float kroot, float dkroot, float eps,
                                            W tiny) {
  if (!(h > tiny)) return 0.f;
  W velg = (hux - u * hx + hvy - v * hy) * ih;
  float s = (float)f_sgn(velg, tiny);
  float k = kc1 * (1.f - s * kroot) - 1.f;
  if (fabs(u) < tiny && fabs(v) < tiny) return 0.f;
  if (k < 0.f) return 0.f;
  return -kc1 * s * dkroot * eps;
}

struct Hfv {
  float u[3], f[3], w[3];
};
// physical flux of one side in direction dir (0 x, 1 y); un = normal velocity
__device__ __forceinline__ Hfv hfv(int dir, float h, float hu, float hv, float un, float k,
                                   float g3, bool wet) {
  Hfv r;
  if (!wet) {
#pragma unroll
    for (int i = 0; i < 3; i++) r.u[i] = r.f[i] = r.w[i] = 0.f;
    return r;
  }
  float a2 = k * h * g3, a = fsqrt(a2);
  r.u[0] = h; r.u[1] = hu; r.u[2] = hv;
  r.f[0] = dir ? hv : hu;
  r.f[1] = hu * un + (dir ? 0.f : 0.5f * a2 * h);
  r.f[2] = hv * un + (dir ? 0.5f * a2 * h : 0.f);
  r.w[0] = un - a; r.w[1] = un; r.w[2] = un + a;
  return r;
}

struct Speeds {
  float sl, sr;
  int mode;  // 0 zero flux, 1 left upwind, 2 right upwind, 3 HLL
};
__device__ __forceinline__ Speeds speeds(const Hfv& l, const Hfv& r) {
  Speeds s;
  if (l.u[0] == 0.f && r.u[0] == 0.f) { s.sl = s.sr = 0.f; s.mode = 0; return s; }
  if (l.u[0] == 0.f) {
    s.sl = fminf(0.f, 2.f * r.w[0] - r.w[1]);
    s.sr = fmaxf(0.f, 2.f * r.w[2] - r.w[1]);
  } else if (r.u[0] == 0.f) {
    s.sl = fminf(0.f, 2.f * l.w[0] - l.w[1]);
    s.sr = fmaxf(0.f, 2.f * l.w[2] - l.w[1]);
  } else {
    s.sl = fminf(0.f, fminf(l.w[0], r.w[0]));
    s.sr = fmaxf(0.f, fmaxf(l.w[2], r.w[2]));
  }
  s.mode = (s.sl >= 0.f) ? 1 : ((s.sr <= 0.f) ? 2 : 3);
  return s;
}

__device__ __forceinline__ void riemann(const Hfv& l, const Hfv& r, float flux[3]) {
  Speeds s = speeds(l, r);
  if (s.mode == 0) { flux[0] = flux[1] = flux[2] = 0.f; return; }
  if (s.mode == 1) { for (int v = 0; v < 3; v++) flux[v] = l.f[v]; return; }
  if (s.mode == 2) { for (int v = 0; v < 3; v++) flux[v] = r.f[v]; return; }
  float inv = rcp(s.sr - s.sl);
#pragma unroll
  for (int v = 0; v < 3; v++)
    flux[v] = (s.sr * l.f[v] - s.sl * r.f[v] + s.sl * s.sr * (r.u[v] - l.u[v])) * inv;
}

}  // namespace m32
}  // namespace tg
