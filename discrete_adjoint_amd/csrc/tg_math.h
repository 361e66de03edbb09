// Per-cell arithmetic of the forward step and its discrete adjoint, as
// device inline functions.  Each function restates one reference routine
// (cited) with the same operand order, so that with -ffp-contract=off and
// IEEE division / square root the FP64 results are bit-identical to the
// reference's.  Used by the FP64 kernels; the mixed-precision kernels use
// their own cheaper formulation.
#pragma once
#include <hip/hip_runtime.h>

namespace tg {

// c_sgn, src/header/geoflow.h:90-100
__device__ __forceinline__ double c_sgn(double z, double tiny) {
  return (z > tiny) ? 1. : ((z < -tiny) ? -1. : 0.);
}
// sgn, src/geoflow/sgn.f:19-31
__device__ __forceinline__ double f_sgn(double z, double tiny) {
  return (fabs(z) < tiny) ? 0. : ((z >= tiny) ? 1. : -1.);
}
// tiny_sgn, src/adjoint/residual.C:153-160
__device__ __forceinline__ double tiny_sgn(double n, double t) {
  return (fabs(n) < t) ? 0. : ((n > t) ? 1. : -1.);
}
// c_dmin1 3-arg, geoflow.h:112-121
__device__ __forceinline__ double dmin3(double a, double b, double c) {
  if (a > b) a = b;
  if (a > c) a = c;
  return a;
}
__device__ __forceinline__ double dabs(double d) { return (d < 0.) ? -d : d; }
// min/max, src/datstr/element2.C:1059-1071
__device__ __forceinline__ double umin(double x, double y) { return (x > y) ? y : x; }
__device__ __forceinline__ double umax(double x, double y) { return (x > y) ? x : y; }

// One limited slope component, forward form (Element::get_slopes,
// element2.C:896-900): divides by the centre spacing.
__device__ __forceinline__ double slope_fwd(double up, double uc, double um, double dx,
                                            double gamma, double tiny) {
  double dp = (up - uc) / dx;
  double dm = (uc - um) / dx;
  double dc = (dp * dx + dm * dx) / (dx + dx);
  return .5 * (c_sgn(dp, tiny) + c_sgn(dm, tiny)) *
         dmin3(gamma * dabs(dp), gamma * dabs(dm), dabs(dc));
}

// Dual form (DualElem::get_slopes_prev, dualmesh.C:1462-1477): multiplies by
// reciprocals.  For the h component also returns hslope_sens{self, minus,
// plus} (dualmesh.C:1479-1532); entries 3,4 (second neighbours) are zero on a
// uniform block.
struct DualSlope {
  double d;
  double s0, s1, s2;
};
__device__ __forceinline__ DualSlope slope_dual(double up, double uc, double um, double dx,
                                                double inv_dx, double inv_s, double gamma,
                                                double tiny, bool want_sens) {
  DualSlope r;
  double dp = (up - uc) * inv_dx;
  double dm = (uc - um) * inv_dx;
  double dc = (dp * dx + dm * dx) * inv_s;
  double mn = dmin3(gamma * dabs(dp), gamma * dabs(dm), dabs(dc));
  double sg = .5 * (c_sgn(dp, tiny) + c_sgn(dm, tiny));
  r.d = sg * mn;
  r.s0 = r.s1 = r.s2 = 0.;
  if (want_sens) {
    if (sg == 0) {
    } else if (mn == gamma * dabs(dp)) {
      r.s0 = -inv_dx; r.s2 = inv_dx;
    } else if (mn == gamma * dabs(dm)) {
      r.s0 = inv_dx; r.s1 = -inv_dx;
    } else {
      r.s0 = inv_dx - inv_dx;
      r.s1 = -inv_dx;
      r.s2 = inv_dx;
      r.s1 *= dx * inv_s;
      r.s2 *= dx * inv_s;
    }
  }
  return r;
}

// gmfggetcoef, src/geoflow/getcoef.f:19-66.  d = (hx, hux, hvx, hy, huy, hvy).
__device__ __forceinline__ double getcoef(double h, double hu, double hv, double hx,
                                          double hux, double hy, double hvy, double kc1,
                                          double kroot, double eps, double tiny) {
  double k;
  if (h > tiny) {
    double vel = hux / h - hu * hx / (h * h) + hvy / h - hv * hy / (h * h);
    k = kc1 * (1. - f_sgn(vel, tiny) * kroot) - 1.;
    if (fabs(hu / h) < tiny && fabs(hv / h) < tiny) k = 1.;
    if (k < 0.) k = 0.;
  } else {
    k = 1.;
  }
  return k * eps;
}

// d kact / d phi_bed (formula branch of getcoef only): no reference
// counterpart, used for the bed-friction gradient.
__device__ __forceinline__ double dgetcoef_dphi(double h, double hu, double hv, double hx,
                                                double hux, double hy, double hvy,
                                                double kc1, double kroot, double dkroot,
                                                double eps, double tiny) {
  if (!(h > tiny)) return 0.;
  double vel = hux / h - hu * hx / (h * h) + hvy / h - hv * hy / (h * h);
  double s = f_sgn(vel, tiny);
  double k = kc1 * (1. - s * kroot) - 1.;
  if (fabs(hu / h) < tiny && fabs(hv / h) < tiny) return 0.;
  if (k < 0.) return 0.;
  return -kc1 * s * dkroot * eps;
}

// eigen, src/geoflow/eigen.f:19-49 -> evalue
__device__ __forceinline__ double eigen(double h, double hu, double hv, double k, double g3,
                                        double tiny) {
  double ex, ey;
  if (h > tiny) {
    double ux = hu / h, uy = hv / h;
    ex = fabs(ux) + sqrt(k * g3 * h);
    ey = fabs(uy) + sqrt(k * g3 * h);
  } else {
    ex = ey = tiny;
  }
  return (ex > ey) ? ex : ey;
}

// State / flux / wave-speed triple of one cell for one direction
// (Element::xdirflux / ydirflux, element2.C:1810-1855 / 1926-1973, dz = 0).
struct Hfv {
  double u[3], f[3], w[3];
};
__device__ __forceinline__ Hfv hfv_fwd(int dir, double h, double hu, double hv, double k,
                                       double g3, double tiny) {
  Hfv r;
  if (h < tiny) {
#pragma unroll
    for (int i = 0; i < 3; i++) r.u[i] = r.f[i] = r.w[i] = 0.;
    return r;
  }
  r.u[0] = h; r.u[1] = hu; r.u[2] = hv;
  if (dir == 0) {
    double vel = hu / h;
    double a = sqrt(k * h * g3);
    r.f[0] = hu;
    r.f[1] = hu * vel + 0.5 * a * a * h;
    r.f[2] = hv * vel;
    r.w[0] = vel - a; r.w[1] = vel; r.w[2] = vel + a;
  } else {
    double a = sqrt(k * h * g3);
    double vel = hv / h;
    r.f[0] = hv;
    r.f[1] = hu * vel;
    r.f[2] = hv * vel + 0.5 * a * a * h;
    r.w[0] = vel - a; r.w[1] = vel; r.w[2] = vel + a;
  }
  return r;
}

// riemannflux (HLL), element2.C:2228-2264
__device__ __forceinline__ void riemann(const Hfv& l, const Hfv& r, double flux[3]) {
  if (l.u[0] == 0. && r.u[0] == 0.) {
    flux[0] = flux[1] = flux[2] = 0.;
    return;
  }
  double sl, sr;
  if (l.u[0] == 0.) {
    sl = umin(0, 2.0 * r.w[0] - r.w[1]);
    sr = umax(0, 2.0 * r.w[2] - r.w[1]);
  } else if (r.u[0] == 0.) {
    sl = umin(0, 2.0 * l.w[0] - l.w[1]);
    sr = umax(0, 2.0 * l.w[2] - l.w[1]);
  } else {
    sl = umin(0, umin(l.w[0], r.w[0]));
    sr = umax(0, umax(l.w[2], r.w[2]));
  }
  if (sl >= 0.0) {
#pragma unroll
    for (int v = 0; v < 3; v++) flux[v] = l.f[v];
  } else if (sr <= 0.0) {
#pragma unroll
    for (int v = 0; v < 3; v++) flux[v] = r.f[v];
  } else {
#pragma unroll
    for (int v = 0; v < 3; v++)
      flux[v] = (sr * l.f[v] - sl * r.f[v] + sl * sr * (r.u[v] - l.u[v])) / (sr - sl);
  }
}

// Source-term switches of update_states (residual.C:200-262).
struct SrcInfo {
  double sgn[4];
  int stop[2];
};

// update_states, src/adjoint/residual.C:162-275.  prev = U^n; f* = face
// fluxes; dUx = (hx, hux, hvx), dUy = (hy, huy, hvy).
__device__ __forceinline__ void update_states(
    const double prev[3], double out[3], const double fxp[3], const double fyp[3],
    const double fxm[3], const double fym[3], double dtdx, double dtdy, double dt,
    double hx, double hvx, double hy, double huy, double k0, double k1, double sin_int,
    double tan_bed, double g0, double g1, double g2, double dg0, double dg1,
    double kact_elem, double ft, double tiny, SrcInfo* si) {
  double vel0, vel1, kact;
  if (prev[0] > tiny) {
    kact = kact_elem;
    vel0 = prev[1] / prev[0];
    vel1 = prev[2] / prev[0];
  } else {
    kact = 0.; vel0 = vel1 = 0.;
  }
#pragma unroll
  for (int i = 0; i < 3; i++)
    out[i] = prev[i] - dtdx * (fxp[i] - fxm[i]) - dtdy * (fyp[i] - fym[i]);
  if (out[0] < 0.) out[0] = 0;
  si->stop[0] = si->stop[1] = 0;
  si->sgn[0] = si->sgn[1] = si->sgn[2] = si->sgn[3] = 0.;
  if (prev[0] > tiny) {
    double h_inv = 1. / prev[0];
    double tmp = h_inv * (huy - vel0 * hy);
    si->sgn[0] = tiny_sgn(tmp, ft);
    si->sgn[2] = tiny_sgn(vel0, ft);
    tmp = h_inv * (hvx - vel1 * hx);
    si->sgn[1] = tiny_sgn(tmp, ft);
    si->sgn[3] = tiny_sgn(vel1, ft);

    double s1 = g0 * prev[0];
    double s2 = si->sgn[0] * prev[0] * kact * (g2 * hy + dg1 * prev[0]) * sin_int;
    double s3 = si->sgn[2] * umax(g2 * prev[0] + vel0 * prev[1] * k0, 0.0) * tan_bed;
    if (s3 == 0. && si->sgn[2] != 0.) si->stop[0] = 1;
    out[1] += dt * (s1 - s2 - s3);

    s1 = g1 * prev[0];
    s2 = si->sgn[1] * prev[0] * kact * (g2 * hx + dg0 * prev[0]) * sin_int;
    s3 = si->sgn[3] * umax(g2 * prev[0] + vel1 * prev[2] * k1, 0.0) * tan_bed;
    if (s3 == 0. && si->sgn[3] != 0.) si->stop[1] = 1;
    out[2] += dt * (s1 - s2 - s3);
  }
}

// Only the switches of update_states (they do not depend on the fluxes).
__device__ __forceinline__ SrcInfo src_switches(double h, double hu, double hv, double hx,
                                                double hvx, double hy, double huy, double k0,
                                                double k1, double tan_bed, double g2,
                                                double ft, double tiny) {
  SrcInfo si;
  si.stop[0] = si.stop[1] = 0;
  si.sgn[0] = si.sgn[1] = si.sgn[2] = si.sgn[3] = 0.;
  if (h > tiny) {
    double vel0 = hu / h, vel1 = hv / h;
    double h_inv = 1. / h;
    double tmp = h_inv * (huy - vel0 * hy);
    si.sgn[0] = tiny_sgn(tmp, ft);
    si.sgn[2] = tiny_sgn(vel0, ft);
    tmp = h_inv * (hvx - vel1 * hx);
    si.sgn[1] = tiny_sgn(tmp, ft);
    si.sgn[3] = tiny_sgn(vel1, ft);
    double s3 = si.sgn[2] * umax(g2 * h + vel0 * hu * k0, 0.0) * tan_bed;
    if (s3 == 0. && si.sgn[2] != 0.) si.stop[0] = 1;
    s3 = si.sgn[3] * umax(g2 * h + vel1 * hv * k1, 0.0) * tan_bed;
    if (s3 == 0. && si.sgn[3] != 0.) si.stop[1] = 1;
  }
  return si;
}

// ---- dual (adjoint) side ----------------------------------------------------

struct M3 {
  double a[3][3];
};
__device__ __forceinline__ M3 zero3() {
  M3 m;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) m.a[i][j] = 0.;
  return m;
}

// Element::dual_xdirflux / dual_ydirflux, element2.C:1858-1923 / 1976-2041
struct DualHfv {
  double u[3], f[3], w[3];
  M3 fj, sj;
};
__device__ __forceinline__ DualHfv dual_hfv(int dir, double h, double hu, double hv, double k,
                                            double g3, double tiny) {
  DualHfv r;
  if (h < tiny) {
#pragma unroll
    for (int i = 0; i < 3; i++) r.u[i] = r.f[i] = r.w[i] = 0.;
    r.fj = zero3(); r.sj = zero3();
    return r;
  }
  r.u[0] = h; r.u[1] = hu; r.u[2] = hv;
  if (dir == 0) {
    double h_inv = 1. / h;
    double vel = hu * h_inv;
    double a = sqrt(k * h * g3);
    r.f[0] = hu; r.f[1] = hu * vel + 0.5 * a * a * h; r.f[2] = hv * vel;
    r.w[0] = vel - a; r.w[1] = vel; r.w[2] = vel + a;
    r.fj.a[0][0] = 0.; r.fj.a[0][1] = 1.; r.fj.a[0][2] = 0.;
    r.fj.a[1][0] = a * a - vel * vel; r.fj.a[1][1] = 2 * vel; r.fj.a[1][2] = 0.;
    r.fj.a[2][0] = -vel * hv * h_inv; r.fj.a[2][1] = hv * h_inv; r.fj.a[2][2] = vel;
    r.sj.a[0][0] = -.5 * a * h_inv - vel * h_inv; r.sj.a[0][1] = h_inv; r.sj.a[0][2] = 0.;
    r.sj.a[1][0] = -vel * h_inv; r.sj.a[1][1] = h_inv; r.sj.a[1][2] = 0.;
    r.sj.a[2][0] = .5 * a * h_inv - vel * h_inv; r.sj.a[2][1] = h_inv; r.sj.a[2][2] = 0.;
  } else {
    double a = sqrt(k * h * g3);
    double h_inv = 1. / h;
    double vel = hv * h_inv;
    r.f[0] = hv; r.f[1] = hu * vel; r.f[2] = hv * vel + 0.5 * a * a * h;
    r.w[0] = vel - a; r.w[1] = vel; r.w[2] = vel + a;
    r.fj.a[0][0] = 0.; r.fj.a[0][1] = 0.; r.fj.a[0][2] = 1.;
    r.fj.a[1][0] = -vel * hu * h_inv; r.fj.a[1][1] = vel; r.fj.a[1][2] = hu * h_inv;
    r.fj.a[2][0] = a * a - vel * vel; r.fj.a[2][1] = 0.; r.fj.a[2][2] = 2 * vel;
    r.sj.a[0][0] = -.5 * a * h_inv - vel * h_inv; r.sj.a[0][1] = 0.; r.sj.a[0][2] = h_inv;
    r.sj.a[1][0] = -vel * h_inv; r.sj.a[1][1] = 0.; r.sj.a[1][2] = h_inv;
    r.sj.a[2][0] = .5 * a * h_inv - vel * h_inv; r.sj.a[2][1] = 0.; r.sj.a[2][2] = h_inv;
  }
  return r;
}

// dual_riemannflux, src/datstr/dualmesh.C:1628-1843: Jacobian of the HLL
// face flux w.r.t. one side only (want_left selects d/dU_l or d/dU_r).
__device__ __forceinline__ M3 dual_riemann_side(const DualHfv& l, const DualHfv& r,
                                                bool want_left) {
  M3 j = zero3();
  if (l.u[0] == 0. && r.u[0] == 0.) return j;
  double sl, sr, dsl_l[3], dsl_r[3], dsr_l[3], dsr_r[3];
  if (l.u[0] == 0.) {
#pragma unroll
    for (int i = 0; i < 3; i++) dsr_l[i] = dsl_l[i] = 0.;
    sl = fmin(0., 2. * r.w[0] - r.w[1]);
#pragma unroll
    for (int i = 0; i < 3; i++) dsl_r[i] = (sl == 0) ? 0. : 2 * r.sj.a[0][i] - r.sj.a[1][i];
    sr = fmax(0., 2. * r.w[2] - r.w[1]);
#pragma unroll
    for (int i = 0; i < 3; i++) dsr_r[i] = (sr == 0) ? 0. : 2 * r.sj.a[2][i] - r.sj.a[1][i];
  } else if (r.u[0] == 0.) {
#pragma unroll
    for (int i = 0; i < 3; i++) dsr_r[i] = dsl_r[i] = 0.;
    sl = fmin(0., 2. * l.w[0] - l.w[1]);
#pragma unroll
    for (int i = 0; i < 3; i++) dsl_l[i] = (sl == 0) ? 0. : 2 * l.sj.a[0][i] - l.sj.a[1][i];
    sr = fmax(0., 2. * l.w[2] - l.w[1]);
#pragma unroll
    for (int i = 0; i < 3; i++) dsr_l[i] = (sr == 0) ? 0. : 2 * l.sj.a[2][i] - l.sj.a[1][i];
  } else {
    sl = fmin(0., fmin(l.w[0], r.w[0]));
    bool sl_zero = (sl == 0), sl_left = (sl == l.w[0]);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      dsl_l[i] = (!sl_zero && sl_left) ? l.sj.a[0][i] : 0.;
      dsl_r[i] = (!sl_zero && !sl_left) ? r.sj.a[0][i] : 0.;
    }
    sr = fmax(0., fmax(l.w[2], r.w[2]));
    bool sr_zero = (sr == 0), sr_left = (sr == l.w[2]);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      dsr_l[i] = (!sr_zero && sr_left) ? l.sj.a[2][i] : 0.;
      dsr_r[i] = (!sr_zero && !sr_left) ? r.sj.a[2][i] : 0.;
    }
  }
  if (sl >= 0.) {
    if (want_left) j = l.fj;
  } else if (sr <= 0.) {
    if (!want_left) j = r.fj;
  } else {
    double inv = 1. / (sr - sl), vd[3], flux[3], dfdsr[3], dfdsl[3];
#pragma unroll
    for (int i = 0; i < 3; i++) vd[i] = r.u[i] - l.u[i];
#pragma unroll
    for (int i = 0; i < 3; i++) flux[i] = (sr * l.f[i] - sl * r.f[i] + sl * sr * vd[i]) * inv;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      dfdsr[i] = (l.f[i] + sl * vd[i] - flux[i]) * inv;
      dfdsl[i] = (-r.f[i] + sr * vd[i] + flux[i]) * inv;
    }
    if (want_left) {
      double jvl = -sr * sl * inv, dFl = sr * inv;
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++)
          j.a[i][k] = dfdsr[i] * dsr_l[k] + dfdsl[i] * dsl_l[k] + dFl * l.fj.a[i][k] +
                      ((i == k) ? jvl : 0.);
    } else {
      double jvr = sr * sl * inv, dFr = -sl * inv;
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++)
          j.a[i][k] = dfdsr[i] * dsr_r[k] + dfdsl[i] * dsl_r[k] + dFr * r.fj.a[i][k] +
                      ((i == k) ? jvr : 0.);
    }
  }
  return j;
}

// Both one-sided Jacobians of one HLL face at once (dual_riemannflux computes
// them together, dualmesh.C:1628-1843): the same expressions as
// dual_riemann_side for each side, the shared intermediates evaluated once.
__device__ __forceinline__ void dual_riemann_both(const DualHfv& l, const DualHfv& r, M3& jl,
                                                  M3& jr) {
  jl = zero3();
  jr = zero3();
  if (l.u[0] == 0. && r.u[0] == 0.) return;
  double sl, sr, dsl_l[3], dsl_r[3], dsr_l[3], dsr_r[3];
  if (l.u[0] == 0.) {
#pragma unroll
    for (int i = 0; i < 3; i++) dsr_l[i] = dsl_l[i] = 0.;
    sl = fmin(0., 2. * r.w[0] - r.w[1]);
#pragma unroll
    for (int i = 0; i < 3; i++) dsl_r[i] = (sl == 0) ? 0. : 2 * r.sj.a[0][i] - r.sj.a[1][i];
    sr = fmax(0., 2. * r.w[2] - r.w[1]);
#pragma unroll
    for (int i = 0; i < 3; i++) dsr_r[i] = (sr == 0) ? 0. : 2 * r.sj.a[2][i] - r.sj.a[1][i];
  } else if (r.u[0] == 0.) {
#pragma unroll
    for (int i = 0; i < 3; i++) dsr_r[i] = dsl_r[i] = 0.;
    sl = fmin(0., 2. * l.w[0] - l.w[1]);
#pragma unroll
    for (int i = 0; i < 3; i++) dsl_l[i] = (sl == 0) ? 0. : 2 * l.sj.a[0][i] - l.sj.a[1][i];
    sr = fmax(0., 2. * l.w[2] - l.w[1]);
#pragma unroll
    for (int i = 0; i < 3; i++) dsr_l[i] = (sr == 0) ? 0. : 2 * l.sj.a[2][i] - l.sj.a[1][i];
  } else {
    sl = fmin(0., fmin(l.w[0], r.w[0]));
    bool sl_zero = (sl == 0), sl_left = (sl == l.w[0]);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      dsl_l[i] = (!sl_zero && sl_left) ? l.sj.a[0][i] : 0.;
      dsl_r[i] = (!sl_zero && !sl_left) ? r.sj.a[0][i] : 0.;
    }
    sr = fmax(0., fmax(l.w[2], r.w[2]));
    bool sr_zero = (sr == 0), sr_left = (sr == l.w[2]);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      dsr_l[i] = (!sr_zero && sr_left) ? l.sj.a[2][i] : 0.;
      dsr_r[i] = (!sr_zero && !sr_left) ? r.sj.a[2][i] : 0.;
    }
  }
  if (sl >= 0.) {
    jl = l.fj;
  } else if (sr <= 0.) {
    jr = r.fj;
  } else {
    double inv = 1. / (sr - sl), vd[3], flux[3], dfdsr[3], dfdsl[3];
#pragma unroll
    for (int i = 0; i < 3; i++) vd[i] = r.u[i] - l.u[i];
#pragma unroll
    for (int i = 0; i < 3; i++) flux[i] = (sr * l.f[i] - sl * r.f[i] + sl * sr * vd[i]) * inv;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      dfdsr[i] = (l.f[i] + sl * vd[i] - flux[i]) * inv;
      dfdsl[i] = (-r.f[i] + sr * vd[i] + flux[i]) * inv;
    }
    const double jvl = -sr * sl * inv, dFl = sr * inv;
    const double jvr = sr * sl * inv, dFr = -sl * inv;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        jl.a[i][k] = dfdsr[i] * dsr_l[k] + dfdsl[i] * dsl_l[k] + dFl * l.fj.a[i][k] +
                     ((i == k) ? jvl : 0.);
        jr.a[i][k] = dfdsr[i] * dsr_r[k] + dfdsl[i] * dsl_r[k] + dFr * r.fj.a[i][k] +
                     ((i == k) ? jvr : 0.);
      }
  }
}

// ---- bed-friction gradient (no reference counterpart) ----------------------

// d(flux)/dk for the forward physical flux of one side
__device__ __forceinline__ void dhfv_dk(int dir, double h, double k, double g3, double tiny,
                                        double dF[3], double dW[3]) {
  dF[0] = dF[1] = dF[2] = dW[0] = dW[1] = dW[2] = 0.;
  if (h < tiny) return;
  double a = sqrt(k * h * g3);
  double da = (a > 0.) ? (h * g3) / (2. * a) : 0.;
  if (dir == 0) dF[1] = 0.5 * h * g3 * h;
  else dF[2] = 0.5 * h * g3 * h;
  dW[0] = -da;
  dW[2] = da;
}

// d(riemannflux)/dk along (dk_l, dk_r) carried in dFl/dWl, dFr/dWr
__device__ __forceinline__ void driemann(const Hfv& l, const Hfv& r, const double dFl[3],
                                         const double dWl[3], const double dFr[3],
                                         const double dWr[3], double out[3]) {
  out[0] = out[1] = out[2] = 0.;
  if (l.u[0] == 0. && r.u[0] == 0.) return;
  double sl, sr, dsl, dsr;
  if (l.u[0] == 0.) {
    double p = 2.0 * r.w[0] - r.w[1], q = 2.0 * r.w[2] - r.w[1];
    sl = umin(0, p); dsl = (0 > p) ? 2.0 * dWr[0] - dWr[1] : 0.;
    sr = umax(0, q); dsr = (0 > q) ? 0. : 2.0 * dWr[2] - dWr[1];
    if (!(sr > 0)) dsr = 0.;
  } else if (r.u[0] == 0.) {
    double p = 2.0 * l.w[0] - l.w[1], q = 2.0 * l.w[2] - l.w[1];
    sl = umin(0, p); dsl = (0 > p) ? 2.0 * dWl[0] - dWl[1] : 0.;
    sr = umax(0, q); dsr = (0 > q) ? 0. : 2.0 * dWl[2] - dWl[1];
    if (!(sr > 0)) dsr = 0.;
  } else {
    double mi = umin(l.w[0], r.w[0]);
    double dmi = (l.w[0] > r.w[0]) ? dWr[0] : dWl[0];
    sl = umin(0, mi); dsl = (0 > mi) ? dmi : 0.;
    double ma = umax(l.w[2], r.w[2]);
    double dma = (l.w[2] > r.w[2]) ? dWl[2] : dWr[2];
    sr = umax(0, ma); dsr = (0 > ma) ? 0. : dma;
    if (!(sr > 0)) dsr = 0.;
  }
  if (sl >= 0.0) {
    for (int v = 0; v < 3; v++) out[v] = dFl[v];
  } else if (sr <= 0.0) {
    for (int v = 0; v < 3; v++) out[v] = dFr[v];
  } else {
    double inv = 1. / (sr - sl);
    for (int v = 0; v < 3; v++) {
      double vd = r.u[v] - l.u[v];
      double f = (sr * l.f[v] - sl * r.f[v] + sl * sr * vd) * inv;
      double dfdsr = (l.f[v] + sl * vd - f) * inv;
      double dfdsl = (-r.f[v] + sr * vd + f) * inv;
      out[v] = dfdsr * dsr + dfdsl * dsl + sr * inv * dFl[v] - sl * inv * dFr[v];
    }
  }
}

}  // namespace tg
