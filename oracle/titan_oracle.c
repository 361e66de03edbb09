/*
 * titan_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never shipped).
 *
 * FP64 CPU restatement of the reference's order-1 forward step and its
 * per-step discrete adjoint on a uniform grid.  The code is organised the way
 * the reference is (per-element slopes, face fluxes stored on edge nodes, a
 * FluxJac container per element, 9 Jacobian blocks per element, adjoint
 * gather through which_neighbor) so that it is an independent restatement of
 * the product's fused GPU kernels.  Expression trees follow the reference
 * operand-for-operand; build with -ffp-contract=off.
 */
#include "titan_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NV 3
#define INIT_NB (-1)

/* Reusable scratch (one buffer per slot, grown on demand): fresh multi-GB
 * allocations per call would spend most of the time in page faults. */
enum { WS_KACT, WS_DU, WS_FX, WS_FY, WS_UN, WS_XL, WS_YL, WS_DUAL, WS_D, WS_K, WS_DK,
       WS_DPHI, WS_N };
static void* ws_ptr[WS_N];
static size_t ws_size[WS_N];
static void* ws(int slot, size_t bytes, int zero) {
  if (bytes > ws_size[slot]) {
    free(ws_ptr[slot]);
    ws_ptr[slot] = malloc(bytes);
    ws_size[slot] = ws_ptr[slot] ? bytes : 0;
  }
  if (zero && ws_ptr[slot]) memset(ws_ptr[slot], 0, bytes);
  return ws_ptr[slot];
}

/* ---- scalar helpers ------------------------------------------------- */

/* c_sgn, src/header/geoflow.h:90-100 (threshold GEOFLOW_TINY) */
static double c_sgn(double z, double tiny) {
  if (z > tiny) return 1.;
  else if (z < -tiny) return -1.;
  return 0.;
}
/* sgn, src/geoflow/sgn.f:19-31 */
static double f_sgn(double z, double tiny) {
  if (fabs(z) < tiny) return 0.;
  else if (z >= tiny) return 1.;
  return -1.;
}
/* tiny_sgn, src/adjoint/residual.C:153-160 */
static double tiny_sgn(double n, double t) {
  if (fabs(n) < t) return 0.;
  else if (n > t) return 1.;
  return -1.;
}
/* c_dmin1 (2 and 3 args), geoflow.h:103-121 */
static double dmin2(double a, double b) { if (a > b) a = b; return a; }
static double dmin3(double a, double b, double c) {
  if (a > b) a = b;
  if (a > c) a = c;
  return a;
}
/* dabs, geoflow.h:143-148 */
static double dabs(double d) { if (d < 0.) d = -d; return d; }
/* min/max, src/datstr/element2.C:1059-1071 */
static double umin(double x, double y) { if (x > y) return y; return x; }
static double umax(double x, double y) { if (x > y) return x; return y; }

/* ---- mesh helpers ---------------------------------------------------- */

static int nbr_of(const or_mesh* m, int c, int d) {
  int i = c % m->nx, j = c / m->nx;
  switch (d) {
    case 0: return (i + 1 < m->nx) ? c + 1 : INIT_NB;
    case 1: return (j + 1 < m->ny) ? c + m->nx : INIT_NB;
    case 2: return (i > 0) ? c - 1 : INIT_NB;
    default: return (j > 0) ? c - m->nx : INIT_NB;
  }
}
static double xc(const or_mesh* m, int c) { return m->x0 + (double)(c % m->nx) * m->dx; }
static double yc(const or_mesh* m, int c) { return m->y0 + (double)(c / m->nx) * m->dy; }
static int on_boundary(const or_mesh* m, int c) {
  for (int d = 0; d < 4; d++)
    if (nbr_of(m, c, d) == INIT_NB) return 1;
  return 0;
}
/* edge-node index of a face: x faces (nx+1)*ny, y faces nx*(ny+1) */
static int xface(const or_mesh* m, int c, int east) {
  int i = c % m->nx, j = c / m->nx;
  return j * (m->nx + 1) + i + east;
}
static int yface(const or_mesh* m, int c, int north) {
  int i = c % m->nx, j = c / m->nx;
  return (j + north) * m->nx + i;
}

/* ---- slopes ----------------------------------------------------------- */

/* Element::get_slopes, src/datstr/element2.C:830-939 (forward: divides). */
static void slopes_fwd(const or_mesh* m, double gamma, double tiny,
                       const double* U[3], int c, double d[6]) {
  if (on_boundary(m, c)) {
    for (int k = 0; k < 6; k++) d[k] = 0.;
    return;
  }
  for (int dir = 0; dir < 2; dir++) {
    int ep = nbr_of(m, c, dir), em = nbr_of(m, c, dir + 2);
    double dxp, dxm;
    if (dir == 0) { dxp = xc(m, ep) - xc(m, c); dxm = xc(m, c) - xc(m, em); }
    else          { dxp = yc(m, ep) - yc(m, c); dxm = yc(m, c) - yc(m, em); }
    for (int j = 0; j < NV; j++) {
      double dp = (U[j][ep] - U[j][c]) / dxp;
      double dm = (U[j][c] - U[j][em]) / dxm;
      double dc = (dp * dxm + dm * dxp) / (dxm + dxp);
      d[j + NV * dir] = .5 * (c_sgn(dp, tiny) + c_sgn(dm, tiny))
                        * dmin3(gamma * dabs(dp), gamma * dabs(dm), dabs(dc));
    }
  }
}

/* DualElem::get_slopes_prev, src/datstr/dualmesh.C:1393-1626 (multiplies by
 * reciprocals; also returns hslope_sens(dir, {self,minus1,plus1,minus2,plus2})). */
static void slopes_dual(const or_mesh* m, double gamma, double tiny,
                        const double* U[3], int c, double d[6], double hs[2][5]) {
  if (on_boundary(m, c)) {
    for (int k = 0; k < 6; k++) d[k] = 0.;
    for (int a = 0; a < 2; a++)
      for (int k = 0; k < 5; k++) hs[a][k] = 0.;
    return;
  }
  for (int dir = 0; dir < 2; dir++) {
    int ep = nbr_of(m, c, dir), em = nbr_of(m, c, dir + 2);
    double dxp, dxm;
    if (dir == 0) { dxp = xc(m, ep) - xc(m, c); dxm = xc(m, c) - xc(m, em); }
    else          { dxp = yc(m, ep) - yc(m, c); dxm = yc(m, c) - yc(m, em); }
    double inv_dxp = 1 / dxp, inv_dxm = 1 / dxm, inv_s = 1 / (dxm + dxp);
    for (int j = 0; j < NV; j++) {
      double dp = (U[j][ep] - U[j][c]) * inv_dxp;
      double dm = (U[j][c] - U[j][em]) * inv_dxm;
      double dc = (dp * dxm + dm * dxp) * inv_s;
      double mn = dmin3(gamma * dabs(dp), gamma * dabs(dm), dabs(dc));
      double sg = .5 * (c_sgn(dp, tiny) + c_sgn(dm, tiny));
      d[j + NV * dir] = sg * mn;
      if (j == 0) {
        double* s = hs[dir];
        if (sg == 0) {
          for (int k = 0; k < 5; k++) s[k] = 0.;
        } else if (mn == gamma * dabs(dp)) {
          s[0] = -inv_dxp; s[1] = 0.; s[2] = inv_dxp; s[3] = 0.; s[4] = 0.;
        } else if (mn == gamma * dabs(dm)) {
          s[0] = inv_dxm; s[1] = -inv_dxm; s[2] = 0.; s[3] = 0.; s[4] = 0.;
        } else {
          s[0] = inv_dxm - inv_dxp; s[1] = -inv_dxm; s[2] = inv_dxp; s[3] = 0.; s[4] = 0.;
          s[1] *= dxp * inv_s;
          s[2] *= dxm * inv_s;
          s[3] *= dxp * inv_s;
          s[4] *= dxm * inv_s;
        }
      }
    }
  }
}

/* ---- earth-pressure coefficient and wave speeds ----------------------- */

/* gmfggetcoef, src/geoflow/getcoef.f:19-66 */
static double getcoef(const or_params* p, double h, double hu, double hv,
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
static double dgetcoef_dphi(const or_params* p, double h, double hu, double hv,
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
static double eigen(const or_params* p, double h, double hu, double hv,
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
static void hfv_fwd(const or_params* p, int dir, double h, double hu, double hv,
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
static void riemann(double l[3][3], double r[3][3], double flux[3]) {
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

static double incrtime(const or_params* p, double time, double dt) {
  /* TimeProps::incrtime, src/header/properties.h:397-410 */
  if (time + dt > p->nextoutput) dt = p->nextoutput - time;
  if (time + dt > p->nextsave) dt = p->nextsave - time;
  if (time + dt > p->maxtime) dt = p->maxtime - time;
  return dt;
}

static double cfl_dt(const or_mesh* m, const or_params* p, const double* U[3],
                     const double* kact) {
  /* get_coef_and_eigen.C:29-180 (single rank: the MPI MIN is the local min) */
  int n = m->nx * m->ny;
  double min_dxdy_ev = 10000000, hmax = 0;
  for (int c = 0; c < n; c++) {
    if (U[0][c] > p->tiny) {
      if (hmax < U[0][c]) hmax = U[0][c];
      double ev = eigen(p, U[0][c], U[1][c], U[2][c], kact[c], m->gz[c]);
      min_dxdy_ev = dmin2(dmin2(m->dx, m->dy) / ev, min_dxdy_ev);
    }
  }
  double dt0 = p->cfl_inner * min_dxdy_ev;
  double dt1 = -0.9 * sqrt(hmax * p->epsilon * p->gscale / 9.8);
  double dt = p->cfl_outer * dmin2(dt0, -dt1);
  if (dt == 0.0) dt = 0.5 * HUGE_VAL;
  return dt;
}

static void all_kact_fwd(const or_mesh* m, const or_params* p, const double* U[3],
                         double* kact, double* dU) {
  int n = m->nx * m->ny;
  for (int c = 0; c < n; c++) {
    slopes_fwd(m, p->gamma, p->tiny, U, c, dU + 6 * c);
    if (U[0][c] > p->tiny)
      kact[c] = getcoef(p, U[0][c], U[1][c], U[2][c], dU + 6 * c);
    else
      kact[c] = p->epsilon;
  }
}

double or_compute_dt(const or_mesh* m, const or_params* p, const double* h,
                     const double* hu, const double* hv, double time) {
  int n = m->nx * m->ny;
  const double* U[3] = {h, hu, hv};
  double* kact = (double*)ws(WS_KACT, sizeof(double) * n, 0);
  double* dU = (double*)ws(WS_DU, sizeof(double) * 6 * n, 0);
  all_kact_fwd(m, p, U, kact, dU);
  return incrtime(p, time, cfl_dt(m, p, U, kact));
}

/* ---- update_states (residual.C:162-275) --------------------------------- */

typedef struct src_info { double sgn[4]; int stop[2]; } src_info;

static void update_states(const or_params* p, const double prev[3], double out[3],
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

/* ---- forward step (step.C:29-275, order 1) ------------------------------- */

int or_step(const or_mesh* m, const or_params* p, double frict_tiny, double* h,
            double* hu, double* hv, double dt_in, double* time, double* dt_out,
            double* mass_out) {
  int n = m->nx * m->ny, nx = m->nx, ny = m->ny;
  const double* U[3] = {h, hu, hv};
  double* kact = (double*)ws(WS_KACT, sizeof(double) * n, 0);
  double* dU = (double*)ws(WS_DU, sizeof(double) * 6 * n, 0);
  double* fx = (double*)ws(WS_FX, sizeof(double) * (size_t)(nx + 1) * ny * NV, 1);
  double* fy = (double*)ws(WS_FY, sizeof(double) * (size_t)nx * (ny + 1) * NV, 1);
  double* Un = (double*)ws(WS_UN, sizeof(double) * 3 * n, 0);
  int* xlist = (int*)ws(WS_XL, sizeof(int) * 2 * n, 0);
  int* ylist = (int*)ws(WS_YL, sizeof(int) * 2 * n, 0);
  int nxl = 0, nyl = 0, rc = 0;

  /* slopes() + get_coef_and_eigen() */
  all_kact_fwd(m, p, U, kact, dU);
  double dt;
  if (dt_in > 0) {
    dt = dt_in;
  } else {
    dt = incrtime(p, *time, cfl_dt(m, p, U, kact));
  }
  *time += dt;
  if (dt_out) *dt_out = dt;

  /* calc_edge_states<Element> (edge_states.C:37-77, element2.C:2744-3072) */
  for (int c = 0; c < n; c++) {
    for (int side = 0; side < 2; side++) {
      int zp = side, zm = side + 2;
      int e = nbr_of(m, c, zp);
      if (e == INIT_NB) {
        if (side) ylist[nyl++] = c; else xlist[nxl++] = c;
      } else {
        double hl[3][3], hr[3][3];
        hfv_fwd(p, side, h[c], hu[c], hv[c], kact[c], m->gz[c], hl);
        hfv_fwd(p, side, h[e], hu[e], hv[e], kact[e], m->gz[e], hr);
        double* dst = side ? fy + NV * yface(m, c, 1) : fx + NV * xface(m, c, 1);
        riemann(hl, hr, dst);
      }
      if (nbr_of(m, c, zm) == INIT_NB) {
        if (side) ylist[nyl++] = c; else xlist[nxl++] = c;
      }
    }
  }
  /* Element::boundary_flux, element2.C:3374-3411 */
  for (int s = 0; s < 2; s++) {
    int cnt = s ? nyl : nxl;
    int* lst = s ? ylist : xlist;
    for (int q = 0; q < cnt; q++) {
      int c = lst[q];
      double* np = s ? fy + NV * yface(m, c, 1) : fx + NV * xface(m, c, 1);
      double* nm = s ? fy + NV * yface(m, c, 0) : fx + NV * xface(m, c, 0);
      if (nbr_of(m, c, s) == INIT_NB)
        for (int v = 0; v < NV; v++) np[v] = nm[v];
      if (nbr_of(m, c, s + 2) == INIT_NB)
        for (int v = 0; v < NV; v++) nm[v] = np[v];
    }
  }

  /* corrector loop (step.C:182-232) -> correct() -> update_states */
  double dtdx = dt / m->dx, dtdy = dt / m->dy, mass = 0.0;
  for (int c = 0; c < n; c++) {
    double prev[3] = {h[c], hu[c], hv[c]}, out[3];
    double g[3] = {m->gx[c], m->gy[c], m->gz[c]};
    double dg[2] = {m->dgx[c], m->dgy[c]};
    double cv[2] = {m->kx[c], m->ky[c]};
    double kc = (prev[0] > p->tiny) ? kact[c] : p->epsilon;
    src_info si;
    update_states(p, prev, out, fx + NV * xface(m, c, 1), fy + NV * yface(m, c, 1),
                  fx + NV * xface(m, c, 0), fy + NV * yface(m, c, 0), dtdx, dtdy, dt,
                  dU + 6 * c, dU + 6 * c + 3, cv, p->sin_int, p->tan_bed, g, dg, kc,
                  frict_tiny, &si);
    for (int v = 0; v < NV; v++) {
      if (!isfinite(out[v])) rc = -5;
      Un[v * n + c] = out[v];
    }
    mass += m->dx * m->dy * out[0];
  }
  memcpy(h, Un, sizeof(double) * n);
  memcpy(hu, Un + n, sizeof(double) * n);
  memcpy(hv, Un + 2 * n, sizeof(double) * n);
  if (mass_out) *mass_out = mass;

  return rc;
}

/* ---- dual fluxes and Jacobians ------------------------------------------- */

typedef double M3[3][3];

/* Element::dual_xdirflux / dual_ydirflux, element2.C:1858-1923, 1976-2041 */
static void dual_hfv(const or_params* p, int dir, double h, double hu, double hv,
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
static void dual_riemann(M3 hl, M3 hr, M3 fjl, M3 fjr, M3 sjl, M3 sjr, M3 jl, M3 jr) {
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

/* Per-element dual data.  On a uniform block every side has one neighbour
 * (neigh_proc[s+4] == -2), so FluxJac index 2 and jacobianMat slots 5..8 are
 * always ZERO_MATRIX (calc_jacobain.C:119, dualmesh.C:1986-1988); they are
 * not stored. */
typedef struct dual_cell {
  double d[6];        /* dual slopes                              */
  double hs[2][5];    /* hslope_sens                              */
  double k;           /* kact from the dual slopes                */
  M3 fjc[2][2][2];    /* FluxJac (side)(dir)(index 0..1)          */
  M3 J[5];            /* jacobianMat slots 0..4                   */
} dual_cell;

static const M3 ZERO3 = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};

/* update_dual_grid (dual_grid_routines.C:15-93) + calc_jacobian
 * (calc_jacobain.C:40-153, set_fluxes_hsens :187-330, calc_jacobian_elem
 * compute_jacobian_elem.C:15-121). */
static dual_cell* build_dual(const or_mesh* m, const or_params* p, double ft,
                             const double* U[3], double dt) {
  int n = m->nx * m->ny;
  dual_cell* dc = (dual_cell*)ws(WS_DUAL, sizeof(dual_cell) * (size_t)n, 1);
  int* xlist = (int*)ws(WS_XL, sizeof(int) * 2 * n, 0);
  int* ylist = (int*)ws(WS_YL, sizeof(int) * 2 * n, 0);
  int nxl = 0, nyl = 0;
  for (int c = 0; c < n; c++) {
    slopes_dual(m, p->gamma, p->tiny, U, c, dc[c].d, dc[c].hs);
    dc[c].k = getcoef(p, U[0][c], U[1][c], U[2][c], dc[c].d);
  }
  /* calc_flux (edge_states.C:79-124) -> DualElem::calc_flux (dualmesh.C:1853-2239) */
  for (int c = 0; c < n; c++) {
    for (int side = 0; side < 2; side++) {
      int zp = side, zm = side + 2;
      int e = nbr_of(m, c, zp);
      if (e == INIT_NB) {
        if (side) ylist[nyl++] = c; else xlist[nxl++] = c;
      } else {
        M3 hl, hr, fjl, fjr, sjl, sjr, jl, jr;
        dual_hfv(p, side, U[0][c], U[1][c], U[2][c], dc[c].k, m->gz[c], hl, fjl, sjl);
        dual_hfv(p, side, U[0][e], U[1][e], U[2][e], dc[e].k, m->gz[e], hr, fjr, sjr);
        dual_riemann(hl, hr, fjl, fjr, sjl, sjr, jl, jr);
        memcpy(dc[c].fjc[side][1][0], jl, sizeof(M3));
        memcpy(dc[c].fjc[side][1][1], jr, sizeof(M3));
        memcpy(dc[e].fjc[side][0][0], jr, sizeof(M3));
        memcpy(dc[e].fjc[side][0][1], jl, sizeof(M3));
      }
      if (nbr_of(m, c, zm) == INIT_NB) {
        if (side) ylist[nyl++] = c; else xlist[nxl++] = c;
      }
    }
  }
  /* DualElem::boundary_flux, dualmesh.C:2241-2279 */
  for (int s = 0; s < 2; s++) {
    int cnt = s ? nyl : nxl;
    int* lst = s ? ylist : xlist;
    for (int q = 0; q < cnt; q++) {
      int c = lst[q];
      if (nbr_of(m, c, s) == INIT_NB)
        for (int i = 0; i < 2; i++) memcpy(dc[c].fjc[s][1][i], dc[c].fjc[s][0][i], sizeof(M3));
      if (nbr_of(m, c, s + 2) == INIT_NB)
        for (int i = 0; i < 2; i++) memcpy(dc[c].fjc[s][0][i], dc[c].fjc[s][1][i], sizeof(M3));
    }
  }
  /* calc_jacobian */
  double dtdx = dt / m->dx, dtdy = dt / m->dy;
  for (int c = 0; c < n; c++) {
    double prev[3] = {U[0][c], U[1][c], U[2][c]}, out[3];
    double g[3] = {m->gx[c], m->gy[c], m->gz[c]};
    double dg[2] = {m->dgx[c], m->dgy[c]};
    double cv[2] = {m->kx[c], m->ky[c]};
    double zf[3] = {0, 0, 0};
    src_info si;
    /* only orgSrcSgn and stop are used; they do not depend on the fluxes */
    update_states(p, prev, out, zf, zf, zf, zf, dtdx, dtdy, dt, dc[c].d, dc[c].d + 3, cv,
                  p->sin_int, p->tan_bed, g, dg, dc[c].k, ft, &si);
    for (int eff = 0; eff < 5; eff++) { /* slots 5..8: neigh_proc == -2, zero */
      M3* J = &dc[c].J[eff];
      int zero = 0;
      if (eff == 0 && prev[0] == 0.) zero = 1;
      if (eff >= 1 && eff <= 4) {
        int nb = nbr_of(m, c, eff - 1);
        if (nb == INIT_NB || U[0][nb] == 0.) zero = 1;
      }
      if (zero) { memcpy(*J, ZERO3, sizeof(M3)); continue; }
      const double (*nxm)[3] = ZERO3, (*pxm)[3] = ZERO3, (*nym)[3] = ZERO3, (*pym)[3] = ZERO3;
      double dh[2] = {0., 0.};
      if (eff == 0) {
        nxm = dc[c].fjc[0][0][0]; pxm = dc[c].fjc[0][1][0];
        nym = dc[c].fjc[1][0][0]; pym = dc[c].fjc[1][1][0];
        dh[0] = dc[c].hs[0][0]; dh[1] = dc[c].hs[1][0];
      } else {
        int dir = (eff - 1) % 4, opos = (dir + 2) % 4;
        int boundary = nbr_of(m, c, opos) == INIT_NB;
        if (dir == 0) {
          dh[0] = dc[c].hs[0][2];
          pxm = dc[c].fjc[0][1][1];
          if (boundary) nxm = dc[c].fjc[0][0][1];
        } else if (dir == 2) {
          dh[0] = dc[c].hs[0][1];
          nxm = dc[c].fjc[0][0][1];
          if (boundary) pxm = dc[c].fjc[0][1][1];
        } else if (dir == 1) {
          dh[1] = dc[c].hs[1][2];
          pym = dc[c].fjc[1][1][1];
          if (boundary) nym = dc[c].fjc[1][0][1];
        } else {
          dh[1] = dc[c].hs[1][1];
          nym = dc[c].fjc[1][0][1];
          if (boundary) pym = dc[c].fjc[1][1][1];
        }
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          (*J)[i][j] = +dtdx * (pxm[i][j] - nxm[i][j]) + dtdy * (pym[i][j] - nym[i][j]);
      if (eff == 0)
        for (int i = 0; i < 3; i++) (*J)[i][i] -= 1.;
      if (prev[0] > p->tiny) {
        double kact = dc[c].k;
        double alpha = p->sin_int * si.sgn[0] * kact;
        double betta = p->sin_int * si.sgn[1] * kact;
        if (eff > 0) {
          if (!si.stop[0]) (*J)[1][0] -= -dt * prev[0] * alpha * g[2] * dh[1];
          if (!si.stop[1]) (*J)[2][0] -= -dt * prev[0] * betta * g[2] * dh[0];
        } else {
          double vx = prev[1] / prev[0], vy = prev[2] / prev[0];
          double vx_sq = vx * vx, vy_sq = vy * vy;
          double tb = p->tan_bed;
          (*J)[1][0] -= dt * (g[0] - alpha * prev[0] * (2 * dg[1] + g[2] * dh[1])
                              - alpha * g[2] * dc[c].d[3]);
          if (si.sgn[2] != 0. && !si.stop[0]) {
            (*J)[1][0] -= -dt * si.sgn[2] * tb * (g[2] - vx_sq * cv[0]);
            (*J)[1][1] -= -dt * si.sgn[2] * tb * (2 * vx * cv[0]);
            (*J)[1][2] -= 0.;
          }
          (*J)[2][0] -= dt * (g[1] - betta * prev[0] * (2 * dg[0] + g[2] * dh[0])
                              - betta * g[2] * dc[c].d[0]);
          if (si.sgn[3] != 0. && !si.stop[1]) {
            (*J)[2][0] -= -dt * si.sgn[3] * tb * (g[2] - vy_sq * cv[1]);
            (*J)[2][1] -= 0.;
            (*J)[2][2] -= -dt * si.sgn[3] * tb * (2 * vy * cv[1]);
          }
        }
      }
    }
  }
  return dc;
}

void or_jacobian(const or_mesh* m, const or_params* p, double frict_tiny,
                 const double* h, const double* hu, const double* hv, double dt,
                 double* Jout) {
  int n = m->nx * m->ny;
  const double* U[3] = {h, hu, hv};
  dual_cell* dc = build_dual(m, p, frict_tiny, U, dt);
  for (int c = 0; c < n; c++)
    for (int s = 0; s < 9; s++)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          Jout[(size_t)c * 81 + s * 9 + i * 3 + j] = (s < 5) ? dc[c].J[s][i][j] : 0.;
}

/* ---- d F / d phi_bed at frozen dt (no reference counterpart) ------------ */

/* d(hfv)/dk for the forward flux: dF[3] and dW[3] (wave speeds). */
static void dhfv_dk(const or_params* p, int dir, double h, double k, double g3,
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
static void driemann(double l[3][3], double r[3][3], const double dFl[3], const double dWl[3],
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
void or_dstep_dphi(const or_mesh* m, const or_params* p, double ft, const double* h,
                   const double* hu, const double* hv, double dt, double* dU) {
  int n = m->nx * m->ny;
  const double* U[3] = {h, hu, hv};
  double* d = (double*)ws(WS_D, sizeof(double) * 6 * n, 0);
  double* k = (double*)ws(WS_K, sizeof(double) * n, 0);
  double* dk = (double*)ws(WS_DK, sizeof(double) * n, 0);
  double hs[2][5];
  for (int c = 0; c < n; c++) {
    slopes_dual(m, p->gamma, p->tiny, U, c, d + 6 * c, hs);
    k[c] = getcoef(p, h[c], hu[c], hv[c], d + 6 * c);
    dk[c] = dgetcoef_dphi(p, h[c], hu[c], hv[c], d + 6 * c);
  }
  double dtdx = dt / m->dx, dtdy = dt / m->dy;
  for (int c = 0; c < n; c++) {
    double flux[4][3], dflux[4][3]; /* E, N, W, S faces */
    for (int f = 0; f < 4; f++) {
      int side = f & 1, plus = f < 2;
      int nb = nbr_of(m, c, f);
      int l = plus ? c : nb, r = plus ? nb : c;
      if (nb == INIT_NB) {
        /* outflow copy of the opposite face, element2.C:3374-3411 */
        int o = nbr_of(m, c, (f + 2) % 4);
        l = plus ? o : c; r = plus ? c : o;
        if (o == INIT_NB) {
          for (int v = 0; v < 3; v++) flux[f][v] = dflux[f][v] = 0.;
          continue;
        }
      }
      double hl[3][3], hr[3][3], dFl[3], dWl[3], dFr[3], dWr[3];
      hfv_fwd(p, side, h[l], hu[l], hv[l], k[l], m->gz[l], hl);
      hfv_fwd(p, side, h[r], hu[r], hv[r], k[r], m->gz[r], hr);
      riemann(hl, hr, flux[f]);
      dhfv_dk(p, side, h[l], k[l], m->gz[l], dFl, dWl);
      dhfv_dk(p, side, h[r], k[r], m->gz[r], dFr, dWr);
      for (int v = 0; v < 3; v++) {
        dFl[v] *= dk[l]; dWl[v] *= dk[l]; dFr[v] *= dk[r]; dWr[v] *= dk[r];
      }
      driemann(hl, hr, dFl, dWl, dFr, dWr, dflux[f]);
    }
    double out[3];
    for (int v = 0; v < 3; v++)
      out[v] = -dtdx * (dflux[0][v] - dflux[2][v]) - dtdy * (dflux[1][v] - dflux[3][v]);
    /* h clamp at zero (residual.C:193-194): a clamped h does not move */
    if (h[c] - dtdx * (flux[0][0] - flux[2][0]) - dtdy * (flux[1][0] - flux[3][0]) < 0.)
      out[0] = 0.;
    if (h[c] > p->tiny) {
      double prev[3] = {h[c], hu[c], hv[c]}, o2[3], zf[3] = {0, 0, 0};
      double g[3] = {m->gx[c], m->gy[c], m->gz[c]}, dg[2] = {m->dgx[c], m->dgy[c]};
      double cv[2] = {m->kx[c], m->ky[c]};
      src_info si;
      update_states(p, prev, o2, zf, zf, zf, zf, dtdx, dtdy, dt, d + 6 * c, d + 6 * c + 3, cv,
                    p->sin_int, p->tan_bed, g, dg, k[c], ft, &si);
      double sec2 = 1. + p->tan_bed * p->tan_bed;
      double vx = prev[1] / prev[0], vy = prev[2] / prev[0];
      const double* dUx = d + 6 * c;
      const double* dUy = d + 6 * c + 3;
      double ds2x = si.sgn[0] * prev[0] * dk[c] * (g[2] * dUy[0] + dg[1] * prev[0]) * p->sin_int;
      double ds3x = si.sgn[2] * umax(g[2] * prev[0] + vx * prev[1] * cv[0], 0.0) * sec2;
      double ds2y = si.sgn[1] * prev[0] * dk[c] * (g[2] * dUx[0] + dg[0] * prev[0]) * p->sin_int;
      double ds3y = si.sgn[3] * umax(g[2] * prev[0] + vy * prev[2] * cv[1], 0.0) * sec2;
      out[1] += dt * (-ds2x - ds3x);
      out[2] += dt * (-ds2y - ds3y);
    }
    dU[c] = out[0];
    dU[n + c] = out[1];
    dU[2 * n + c] = out[2];
  }
}

/* ---- adjoint step (calc_adjoint.C:35-167) -------------------------------- */

void or_adjoint_step(const or_mesh* m, const or_params* p, double frict_tiny,
                     const double* h, const double* hu, const double* hv, double dt,
                     const double* lam_next, const double* g, double* lam_out,
                     double* dJdphi) {
  int n = m->nx * m->ny;
  const double* U[3] = {h, hu, hv};
  dual_cell* dc = build_dual(m, p, frict_tiny, U, dt);
  for (int c = 0; c < n; c++) {
    double adj[3] = {0., 0., 0.};
    /* effelement 1..4 visit neighbor[0..3]; generated grids have
     * positive_x_side = 1 (createfunky.C:310-318 node order,
     * element2.C:694-716), so absolute sides 0..3 are S, E, N, W. */
    static const int side_dir[4] = {3, 0, 1, 2};
    for (int eff = 0; eff < 5; eff++) {
      int src = c, slot = 0;
      if (eff > 0) {
        int d = side_dir[eff - 1];
        src = nbr_of(m, c, d);
        if (src == INIT_NB) continue;
        slot = ((d + 2) % 4) + 1; /* which_neighbor(c) + 1 in the neighbour's slots */
      }
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++)
          adj[k] += lam_next[l * n + src] * dc[src].J[slot][l][k];
    }
    for (int j = 0; j < 3; j++) {
      double gj = g ? g[j * n + c] : 0.;
      double v = gj - adj[j];
      if (fabs(v) < 1e-16) v = 0.;
      lam_out[j * n + c] = v;
    }
  }
  if (dJdphi) {
    double* dU = (double*)ws(WS_DPHI, sizeof(double) * 3 * n, 0);
    or_dstep_dphi(m, p, frict_tiny, h, hu, hv, dt, dU);
    double acc = 0.;
    for (int c = 0; c < n; c++)
      for (int v = 0; v < 3; v++) acc += lam_next[v * n + c] * dU[v * n + c];
    *dJdphi += acc;
  }
}
