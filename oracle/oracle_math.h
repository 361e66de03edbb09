/*
 * oracle_math.h -- TEST INFRASTRUCTURE ONLY (the checker, never shipped).
 *
 * Per-cell arithmetic shared by the uniform-block and the leaf-table CPU
 * restatements: thresholds, getcoef/eigen, physical flux and HLL,
 * update_states, the dual flux / HLL Jacobians and the d/dkact terms of the
 * bed-friction gradient.  Expression trees follow the reference operand for
 * operand; build with -ffp-contract=off.
 */
#ifndef ORACLE_MATH_H
#define ORACLE_MATH_H

#include <math.h>
#include <string.h>

#include "titan_oracle.h"

#define NV 3

/* ---- per-cell material ------------------------------------------------ */

/* The parameters of cell c: p with the bed friction of the cell's material
 * (get_coef_and_eigen.C:112-114 and corrector.C:108 read
 * bedfrict[material]); p itself when the mesh has one material. */
static inline const or_params* cell_params(const or_materials* mm, const or_params* p, int c,
                                           or_params* tmp) {
  if (!mm->mat) return p;
  const int k = mm->mat[c];
  *tmp = *p;
  tmp->tan_bed = mm->tan_bed[k];
  tmp->kact_root = mm->kact_root[k];
  tmp->dkroot_dphi = mm->dkroot[k];
  return tmp;
}

/* ---- scalar helpers ------------------------------------------------- */

/* c_sgn, src/header/geoflow.h:90-100 (threshold GEOFLOW_TINY) */
static inline double c_sgn(double z, double tiny) {
  if (z > tiny) return 1.;
  else if (z < -tiny) return -1.;
  return 0.;
}
/* sgn, src/geoflow/sgn.f:19-31 */
static inline double f_sgn(double z, double tiny) {
  if (fabs(z) < tiny) return 0.;
  else if (z >= tiny) return 1.;
  return -1.;
}
/* tiny_sgn, src/adjoint/residual.C:153-160 */
static inline double tiny_sgn(double n, double t) {
  if (fabs(n) < t) return 0.;
  else if (n > t) return 1.;
  return -1.;
}
/* c_dmin1 (2 and 3 args), geoflow.h:103-121 */
static inline double dmin2(double a, double b) { if (a > b) a = b; return a; }
static inline double dmin3(double a, double b, double c) {
  if (a > b) a = b;
  if (a > c) a = c;
  return a;
}
/* dabs, geoflow.h:143-148 */
static inline double dabs(double d) { if (d < 0.) d = -d; return d; }
/* min/max, src/datstr/element2.C:1059-1071 */
static inline double umin(double x, double y) { if (x > y) return y; return x; }
static inline double umax(double x, double y) { if (x > y) return x; return y; }


/* ---- earth-pressure coefficient and wave speeds ----------------------- */

/* gmfggetcoef, src/geoflow/getcoef.f:19-66 */
static inline double getcoef(const or_params* p, double h, double hu, double hv,
                      const double d[6]) {
  double k;
  if (h > p->tiny) {
    double vel = d[1] / h - hu * d[0] / (h * h) + d[5] / h - hv * d[3] / (h * h);
    k = p->kact_c1 * (1. - f_sgn(vel, p->tiny) * p->kact_root) - 1.;
    if (fabs(hu / h) < p->tiny && fabs(hv / h) < p->tiny) k = 1.;
    if (k < 0.) k = 0.;
  } else {
    k = 1.;
  }
  return k * p->epsilon;
}

/* d kact / d phi_bed along the formula branch of getcoef (0 elsewhere). */
static inline double dgetcoef_dphi(const or_params* p, double h, double hu, double hv,
                            const double d[6]) {
  if (!(h > p->tiny)) return 0.;
  double vel = d[1] / h - hu * d[0] / (h * h) + d[5] / h - hv * d[3] / (h * h);
  double s = f_sgn(vel, p->tiny);
  double k = p->kact_c1 * (1. - s * p->kact_root) - 1.;
  if (fabs(hu / h) < p->tiny && fabs(hv / h) < p->tiny) return 0.;
  if (k < 0.) return 0.;
  return -p->kact_c1 * s * p->dkroot_dphi * p->epsilon;
}

/* eigen, src/geoflow/eigen.f:19-49, returns evalue */
static inline double eigen(const or_params* p, double h, double hu, double hv,
                    double k, double g3) {
  double ex, ey;
  if (h > p->tiny) {
    double ux = hu / h, uy = hv / h;
    ex = fabs(ux) + sqrt(k * g3 * h);
    ey = fabs(uy) + sqrt(k * g3 * h);
  } else {
    ex = ey = p->tiny;
  }
  return (ex > ey) ? ex : ey;
}

/* ---- physical flux and HLL -------------------------------------------- */

/* hfv[3][3]: row 0 state, row 1 flux, row 2 wave speeds.
 * Element::xdirflux / ydirflux, element2.C:1810-1855, 1926-1973 (order 1) */
static inline void hfv_fwd(const or_params* p, int dir, double h, double hu, double hv,
                    double k, double g3, double f[3][3]) {
  if (h < p->tiny) {
    memset(f, 0, sizeof(double) * 9);
    return;
  }
  f[0][0] = h; f[0][1] = hu; f[0][2] = hv;
  if (dir == 0) {
    double vel = f[0][1] / f[0][0];
    double a = sqrt(k * f[0][0] * g3);
    f[1][0] = f[0][1];
    f[1][1] = f[0][1] * vel + 0.5 * a * a * f[0][0];
    f[1][2] = f[0][2] * vel;
    f[2][0] = vel - a; f[2][1] = vel; f[2][2] = vel + a;
  } else {
    double a = sqrt(k * f[0][0] * g3);
    double vel = f[0][2] / f[0][0];
    f[1][0] = f[0][2];
    f[1][1] = f[0][1] * vel;
    f[1][2] = f[0][2] * vel + 0.5 * a * a * f[0][0];
    f[2][0] = vel - a; f[2][1] = vel; f[2][2] = vel + a;
  }
}

/* riemannflux, element2.C:2228-2264 */
static inline void riemann(double l[3][3], double r[3][3], double flux[3]) {
  if (l[0][0] == 0. && r[0][0] == 0.) {
    flux[0] = flux[1] = flux[2] = 0.;
    return;
  }
  double sl, sr;
  if (l[0][0] == 0.) {
    sl = umin(0, 2.0 * r[2][0] - r[2][1]);
    sr = umax(0, 2.0 * r[2][2] - r[2][1]);
  } else if (r[0][0] == 0.) {
    sl = umin(0, 2.0 * l[2][0] - l[2][1]);
    sr = umax(0, 2.0 * l[2][2] - l[2][1]);
  } else {
    sl = umin(0, umin(l[2][0], r[2][0]));
    sr = umax(0, umax(l[2][2], r[2][2]));
  }
  if (sl >= 0.0)
    for (int v = 0; v < NV; v++) flux[v] = l[1][v];
  else if (sr <= 0.0)
    for (int v = 0; v < NV; v++) flux[v] = r[1][v];
  else
    for (int v = 0; v < NV; v++)
      flux[v] = (sr * l[1][v] - sl * r[1][v] + sl * sr * (r[0][v] - l[0][v])) / (sr - sl);
}


/* ---- CFL ---------------------------------------------------------------- */

static inline double incrtime(const or_params* p, double time, double dt) {
  /* TimeProps::incrtime, src/header/properties.h:397-410 */
  if (time + dt > p->nextoutput) dt = p->nextoutput - time;
  if (time + dt > p->nextsave) dt = p->nextsave - time;
  if (time + dt > p->maxtime) dt = p->maxtime - time;
  return dt;
}


/* ---- update_states (residual.C:162-275) --------------------------------- */

typedef struct src_info { double sgn[4]; int stop[2]; } src_info;

static inline void update_states(const or_params* p, const double prev[3], double out[3],
                          const double fxp[3], const double fyp[3],
                          const double fxm[3], const double fym[3],
                          double dtdx, double dtdy, double dt, const double* dUx,
                          const double* dUy, const double curv[2], double sin_int,
                          double tan_bed, const double g[3], const double dg[2],
                          double kact_elem, double ft, src_info* si) {
  double vel[2], kact;
  if (prev[0] > p->tiny) {
    kact = kact_elem;
    vel[0] = prev[1] / prev[0];
    vel[1] = prev[2] / prev[0];
  } else {
    kact = 0.;
    vel[0] = vel[1] = 0.;
  }
  for (int i = 0; i < NV; i++)
    out[i] = prev[i] - dtdx * (fxp[i] - fxm[i]) - dtdy * (fyp[i] - fym[i]);
  if (out[0] < 0.) out[0] = 0;
  si->stop[0] = si->stop[1] = 0;
  si->sgn[0] = si->sgn[1] = si->sgn[2] = si->sgn[3] = 0.;
  if (prev[0] > p->tiny) {
    double h_inv = 1. / prev[0];
    double tmp = h_inv * (dUy[1] - vel[0] * dUy[0]);
    si->sgn[0] = tiny_sgn(tmp, ft);
    si->sgn[2] = tiny_sgn(vel[0], ft);
    tmp = h_inv * (dUx[2] - vel[1] * dUx[0]);
    si->sgn[1] = tiny_sgn(tmp, ft);
    si->sgn[3] = tiny_sgn(vel[1], ft);

    double s1 = g[0] * prev[0];
    double s2 = si->sgn[0] * prev[0] * kact * (g[2] * dUy[0] + dg[1] * prev[0]) * sin_int;
    double s3 = si->sgn[2] * umax(g[2] * prev[0] + vel[0] * prev[1] * curv[0], 0.0) * tan_bed;
    if (s3 == 0. && si->sgn[2]) si->stop[0] = 1;
    out[1] += dt * (s1 - s2 - s3);

    s1 = g[1] * prev[0];
    s2 = si->sgn[1] * prev[0] * kact * (g[2] * dUx[0] + dg[0] * prev[0]) * sin_int;
    s3 = si->sgn[3] * umax(g[2] * prev[0] + vel[1] * prev[2] * curv[1], 0.0) * tan_bed;
    if (s3 == 0. && si->sgn[3]) si->stop[1] = 1;
    out[2] += dt * (s1 - s2 - s3);
  }
}


/* ---- dual fluxes and Jacobians ------------------------------------------- */

typedef double M3[3][3];

/* Element::dual_xdirflux / dual_ydirflux, element2.C:1858-1923, 1976-2041 */
static inline void dual_hfv(const or_params* p, int dir, double h, double hu, double hv,
                     double k, double g3, M3 hfv, M3 fj, M3 sj) {
  if (h < p->tiny) {
    memset(hfv, 0, sizeof(M3)); memset(fj, 0, sizeof(M3)); memset(sj, 0, sizeof(M3));
    return;
  }
  hfv[0][0] = h; hfv[0][1] = hu; hfv[0][2] = hv;
  if (dir == 0) {
    double h_inv = 1. / hfv[0][0];
    double vel = hfv[0][1] * h_inv;
    double a = sqrt(k * hfv[0][0] * g3);
    hfv[1][0] = hfv[0][1];
    hfv[1][1] = hfv[0][1] * vel + 0.5 * a * a * hfv[0][0];
    hfv[1][2] = hfv[0][2] * vel;
    hfv[2][0] = vel - a; hfv[2][1] = vel; hfv[2][2] = vel + a;
    fj[0][0] = 0.; fj[0][1] = 1.; fj[0][2] = 0.;
    fj[1][0] = a * a - vel * vel; fj[1][1] = 2 * vel; fj[1][2] = 0.;
    fj[2][0] = -vel * hfv[0][2] * h_inv; fj[2][1] = hfv[0][2] * h_inv; fj[2][2] = vel;
    sj[0][0] = -.5 * a * h_inv - vel * h_inv; sj[0][1] = h_inv; sj[0][2] = 0.;
    sj[1][0] = -vel * h_inv; sj[1][1] = h_inv; sj[1][2] = 0.;
    sj[2][0] = .5 * a * h_inv - vel * h_inv; sj[2][1] = h_inv; sj[2][2] = 0.;
  } else {
    double a = sqrt(k * hfv[0][0] * g3);
    double h_inv = 1. / hfv[0][0];
    double vel = hfv[0][2] * h_inv;
    hfv[1][0] = hfv[0][2];
    hfv[1][1] = hfv[0][1] * vel;
    hfv[1][2] = hfv[0][2] * vel + 0.5 * a * a * hfv[0][0];
    hfv[2][0] = vel - a; hfv[2][1] = vel; hfv[2][2] = vel + a;
    fj[0][0] = 0.; fj[0][1] = 0.; fj[0][2] = 1.;
    fj[1][0] = -vel * hfv[0][1] * h_inv; fj[1][1] = vel; fj[1][2] = hfv[0][1] * h_inv;
    fj[2][0] = a * a - vel * vel; fj[2][1] = 0.; fj[2][2] = 2 * vel;
    sj[0][0] = -.5 * a * h_inv - vel * h_inv; sj[0][1] = 0.; sj[0][2] = h_inv;
    sj[1][0] = -vel * h_inv; sj[1][1] = 0.; sj[1][2] = h_inv;
    sj[2][0] = .5 * a * h_inv - vel * h_inv; sj[2][1] = 0.; sj[2][2] = h_inv;
  }
}

/* dual_riemannflux, src/datstr/dualmesh.C:1628-1843 (flux values dropped: the
 * adjoint only needs the Jacobians). */
static inline void dual_riemann(M3 hl, M3 hr, M3 fjl, M3 fjr, M3 sjl, M3 sjr, M3 jl, M3 jr) {
  if (hl[0][0] == 0. && hr[0][0] == 0.) {
    memset(jl, 0, sizeof(M3)); memset(jr, 0, sizeof(M3));
    return;
  }
  double sl, sr, dsl_l[3], dsl_r[3], dsr_l[3], dsr_r[3];
  if (hl[0][0] == 0.) {
    for (int i = 0; i < 3; i++) dsr_l[i] = dsl_l[i] = 0.;
    sl = fmin(0., 2. * hr[2][0] - hr[2][1]);
    if (sl == 0) { for (int i = 0; i < 3; i++) dsl_r[i] = 0.; }
    else for (int i = 0; i < 3; i++) dsl_r[i] = 2 * sjr[0][i] - sjr[1][i];
    sr = fmax(0., 2. * hr[2][2] - hr[2][1]);
    if (sr == 0) { for (int i = 0; i < 3; i++) dsr_r[i] = 0.; }
    else for (int i = 0; i < 3; i++) dsr_r[i] = 2 * sjr[2][i] - sjr[1][i];
  } else if (hr[0][0] == 0.) {
    for (int i = 0; i < 3; i++) dsr_r[i] = dsl_r[i] = 0.;
    sl = fmin(0., 2. * hl[2][0] - hl[2][1]);
    if (sl == 0) { for (int i = 0; i < 3; i++) dsl_l[i] = 0.; }
    else for (int i = 0; i < 3; i++) dsl_l[i] = 2 * sjl[0][i] - sjl[1][i];
    sr = fmax(0., 2. * hl[2][2] - hl[2][1]);
    if (sr == 0) { for (int i = 0; i < 3; i++) dsr_l[i] = 0.; }
    else for (int i = 0; i < 3; i++) dsr_l[i] = 2 * sjl[2][i] - sjl[1][i];
  } else {
    sl = fmin(0., fmin(hl[2][0], hr[2][0]));
    if (sl == 0) {
      for (int i = 0; i < 3; i++) dsl_l[i] = dsl_r[i] = 0.;
    } else if (sl == hl[2][0]) {
      for (int i = 0; i < 3; i++) { dsl_r[i] = 0.; dsl_l[i] = sjl[0][i]; }
    } else {
      for (int i = 0; i < 3; i++) { dsl_l[i] = 0.; dsl_r[i] = sjr[0][i]; }
    }
    sr = fmax(0., fmax(hl[2][2], hr[2][2]));
    if (sr == 0) {
      for (int i = 0; i < 3; i++) dsr_l[i] = dsr_r[i] = 0.;
    } else if (sr == hl[2][2]) {
      for (int i = 0; i < 3; i++) { dsr_r[i] = 0.; dsr_l[i] = sjl[2][i]; }
    } else {
      for (int i = 0; i < 3; i++) { dsr_l[i] = 0.; dsr_r[i] = sjr[2][i]; }
    }
  }
  if (sl >= 0.) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) { jr[i][j] = 0.; jl[i][j] = fjl[i][j]; }
  } else if (sr <= 0.) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) { jr[i][j] = fjr[i][j]; jl[i][j] = 0.; }
  } else {
    double inv = 1. / (sr - sl), vd[3], flux[3], dfdsr[3], dfdsl[3];
    for (int i = 0; i < 3; i++) vd[i] = hr[0][i] - hl[0][i];
    for (int i = 0; i < 3; i++)
      flux[i] = (sr * hl[1][i] - sl * hr[1][i] + sl * sr * vd[i]) * inv;
    for (int i = 0; i < 3; i++) {
      dfdsr[i] = (hl[1][i] + sl * vd[i] - flux[i]) * inv;
      dfdsl[i] = (-hr[1][i] + sr * vd[i] + flux[i]) * inv;
    }
    double jvr = sr * sl * inv, jvl = -sr * sl * inv;
    double dFr = -sl * inv, dFl = sr * inv;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double vr = (i == j) ? jvr : 0., vl = (i == j) ? jvl : 0.;
        jr[i][j] = dfdsr[i] * dsr_r[j] + dfdsl[i] * dsl_r[j] + dFr * fjr[i][j] + vr;
        jl[i][j] = dfdsr[i] * dsr_l[j] + dfdsl[i] * dsl_l[j] + dFl * fjl[i][j] + vl;
      }
  }
}


static const M3 ZERO3 = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};


/* ---- d F / d phi_bed at frozen dt (no reference counterpart) ------------ */

/* d(hfv)/dk for the forward flux: dF[3] and dW[3] (wave speeds). */
static inline void dhfv_dk(const or_params* p, int dir, double h, double k, double g3,
                    double dF[3], double dW[3]) {
  dF[0] = dF[1] = dF[2] = dW[0] = dW[1] = dW[2] = 0.;
  if (h < p->tiny) return;
  double a = sqrt(k * h * g3);
  double da = (a > 0.) ? (h * g3) / (2. * a) : 0.;
  if (dir == 0) dF[1] = 0.5 * h * g3 * h;
  else dF[2] = 0.5 * h * g3 * h;
  dW[0] = -da; dW[2] = da;
}

/* derivative of riemannflux w.r.t. (k_l, k_r) times (dk_l, dk_r) */
static inline void driemann(double l[3][3], double r[3][3], const double dFl[3], const double dWl[3],
                     const double dFr[3], const double dWr[3], double out[3]) {
  out[0] = out[1] = out[2] = 0.;
  if (l[0][0] == 0. && r[0][0] == 0.) return;
  double sl, sr, dsl, dsr;
  if (l[0][0] == 0.) {
    double a = 2.0 * r[2][0] - r[2][1], b = 2.0 * r[2][2] - r[2][1];
    sl = umin(0, a); dsl = (0 > a) ? 2.0 * dWr[0] - dWr[1] : 0.;
    sr = umax(0, b); dsr = (0 > b) ? 0. : 2.0 * dWr[2] - dWr[1];
    if (!(sr > 0)) dsr = 0.;
  } else if (r[0][0] == 0.) {
    double a = 2.0 * l[2][0] - l[2][1], b = 2.0 * l[2][2] - l[2][1];
    sl = umin(0, a); dsl = (0 > a) ? 2.0 * dWl[0] - dWl[1] : 0.;
    sr = umax(0, b); dsr = (0 > b) ? 0. : 2.0 * dWl[2] - dWl[1];
    if (!(sr > 0)) dsr = 0.;
  } else {
    double mi = umin(l[2][0], r[2][0]);
    double dmi = (l[2][0] > r[2][0]) ? dWr[0] : dWl[0];
    sl = umin(0, mi); dsl = (0 > mi) ? dmi : 0.;
    double ma = umax(l[2][2], r[2][2]);
    double dma = (l[2][2] > r[2][2]) ? dWl[2] : dWr[2];
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
      double vd = r[0][v] - l[0][v];
      double f = (sr * l[1][v] - sl * r[1][v] + sl * sr * vd) * inv;
      double dfdsr = (l[1][v] + sl * vd - f) * inv;
      double dfdsl = (-r[1][v] + sr * vd + f) * inv;
      out[v] = dfdsr * dsr + dfdsl * dsl + sr * inv * dFl[v] - sl * inv * dFr[v];
    }
  }
}

/* dU^{s+1}/dphi_bed per cell.  kact is the one the adjoint sweep holds
 * (from the dual slopes, dual_grid_routines.C:68-84); the corrector terms
 * follow update_states (residual.C:162-275). */

#endif
