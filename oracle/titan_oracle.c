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
#include "oracle_math.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

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
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; c++) {
    slopes_fwd(m, p->gamma, p->tiny, U, c, dU + 6 * c);
    if (U[0][c] > p->tiny)
    {
      or_params t;
      kact[c] = getcoef(cell_params(&m->mats, p, c, &t), U[0][c], U[1][c], U[2][c], dU + 6 * c);
    }
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

/* The forward slopes and kact of every cell exactly as or_step computes them
 * (slopes() then the get_coef_and_eigen.C:29-115 per-cell coefficient);
 * kact [n], dU [6n] (per cell dh/dx, dhu/dx, dhv/dx, dh/dy, dhu/dy, dhv/dy). */
void or_kact_slopes(const or_mesh* m, const or_params* p, const double* h,
                    const double* hu, const double* hv, double* kact, double* dU) {
  const double* U[3] = {h, hu, hv};
  all_kact_fwd(m, p, U, kact, dU);
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
  if (!isnan(dt_in)) { /* a replayed dt, any value */
    dt = dt_in;
  } else {
    dt = incrtime(p, *time, cfl_dt(m, p, U, kact));
  }
  *time += dt;
  if (dt_out) *dt_out = dt;

  /* calc_edge_states<Element> (edge_states.C:37-77, element2.C:2744-3072):
   * every cell writes the fluxes of its + faces (a face has one writer) */
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; c++) {
    for (int side = 0; side < 2; side++) {
      int e = nbr_of(m, c, side);
      if (e != INIT_NB) {
        double hl[3][3], hr[3][3];
        hfv_fwd(p, side, h[c], hu[c], hv[c], kact[c], m->gz[c], hl);
        hfv_fwd(p, side, h[e], hu[e], hv[e], kact[e], m->gz[e], hr);
        double* dst = side ? fy + NV * yface(m, c, 1) : fx + NV * xface(m, c, 1);
        riemann(hl, hr, dst);
      }
    }
  }
  /* the x / y boundary lists of edge_states.C:37-77, in traversal order */
  for (int c = 0; c < n; c++)
    for (int side = 0; side < 2; side++) {
      if (nbr_of(m, c, side) == INIT_NB) {
        if (side) ylist[nyl++] = c; else xlist[nxl++] = c;
      }
      if (nbr_of(m, c, side + 2) == INIT_NB) {
        if (side) ylist[nyl++] = c; else xlist[nxl++] = c;
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
#pragma omp parallel for schedule(static) reduction(+ : mass) reduction(min : rc)
  for (int c = 0; c < n; c++) {
    double prev[3] = {h[c], hu[c], hv[c]}, out[3];
    double g[3] = {m->gx[c], m->gy[c], m->gz[c]};
    double dg[2] = {m->dgx[c], m->dgy[c]};
    double cv[2] = {m->kx[c], m->ky[c]};
    double kc = (prev[0] > p->tiny) ? kact[c] : p->epsilon;
    src_info si;
    or_params pt;
    update_states(p, prev, out, fx + NV * xface(m, c, 1), fy + NV * yface(m, c, 1),
                  fx + NV * xface(m, c, 0), fy + NV * yface(m, c, 0), dtdx, dtdy, dt,
                  dU + 6 * c, dU + 6 * c + 3, cv, p->sin_int,
                  cell_params(&m->mats, p, c, &pt)->tan_bed, g, dg, kc,
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
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; c++) {
    slopes_dual(m, p->gamma, p->tiny, U, c, dc[c].d, dc[c].hs);
    or_params t;
    dc[c].k = getcoef(cell_params(&m->mats, p, c, &t), U[0][c], U[1][c], U[2][c], dc[c].d);
  }
  /* calc_flux (edge_states.C:79-124) -> DualElem::calc_flux (dualmesh.C:1853-2239):
   * the face between c and its + neighbour e is written by c alone */
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; c++) {
    for (int side = 0; side < 2; side++) {
      int e = nbr_of(m, c, side);
      if (e != INIT_NB) {
        M3 hl, hr, fjl, fjr, sjl, sjr, jl, jr;
        dual_hfv(p, side, U[0][c], U[1][c], U[2][c], dc[c].k, m->gz[c], hl, fjl, sjl);
        dual_hfv(p, side, U[0][e], U[1][e], U[2][e], dc[e].k, m->gz[e], hr, fjr, sjr);
        dual_riemann(hl, hr, fjl, fjr, sjl, sjr, jl, jr);
        memcpy(dc[c].fjc[side][1][0], jl, sizeof(M3));
        memcpy(dc[c].fjc[side][1][1], jr, sizeof(M3));
        memcpy(dc[e].fjc[side][0][0], jr, sizeof(M3));
        memcpy(dc[e].fjc[side][0][1], jl, sizeof(M3));
      }
    }
  }
  for (int c = 0; c < n; c++)
    for (int side = 0; side < 2; side++) {
      if (nbr_of(m, c, side) == INIT_NB) {
        if (side) ylist[nyl++] = c; else xlist[nxl++] = c;
      }
      if (nbr_of(m, c, side + 2) == INIT_NB) {
        if (side) ylist[nyl++] = c; else xlist[nxl++] = c;
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
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; c++) {
    double prev[3] = {U[0][c], U[1][c], U[2][c]}, out[3];
    double g[3] = {m->gx[c], m->gy[c], m->gz[c]};
    double dg[2] = {m->dgx[c], m->dgy[c]};
    double cv[2] = {m->kx[c], m->ky[c]};
    double zf[3] = {0, 0, 0};
    src_info si;
    or_params pt;
    const double tan_bed = cell_params(&m->mats, p, c, &pt)->tan_bed;
    /* only orgSrcSgn and stop are used; they do not depend on the fluxes */
    update_states(p, prev, out, zf, zf, zf, zf, dtdx, dtdy, dt, dc[c].d, dc[c].d + 3, cv,
                  p->sin_int, tan_bed, g, dg, dc[c].k, ft, &si);
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
          double tb = tan_bed;
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

/* Where the reference's adjoint approximations act, per cell (test
 * infrastructure: classifies analytic-vs-FD Jacobian mismatches):
 * sg[2n]   the h-slope limiter's sign factor .5 (c_sgn(dp) + c_sgn(dm)) in x
 *          and y as slopes_dual forms it (0 on the boundary ring); its
 *          hslope_sens ignores the factor, so it is exact only for |sg| = 1
 *          or 0 (element2.C:886-897, dualmesh.C:1393-1626);
 * stop[2n] update_states' Coulomb stop flags: the neighbour blocks drop the
 *          internal-friction slope term under stop (calc_jacobain.C). */
void or_switches(const or_mesh* m, const or_params* p, double ft, const double* h,
                 const double* hu, const double* hv, double dt, double* sg, int* stop) {
  int n = m->nx * m->ny;
  const double* U[3] = {h, hu, hv};
  dual_cell* dc = build_dual(m, p, ft, U, dt);
  for (int c = 0; c < n; c++) {
    sg[c] = sg[n + c] = 0.;
    if (!on_boundary(m, c))
      for (int dir = 0; dir < 2; dir++) {
        int ep = nbr_of(m, c, dir), em = nbr_of(m, c, dir + 2);
        double dxp = dir ? yc(m, ep) - yc(m, c) : xc(m, ep) - xc(m, c);
        double dxm = dir ? yc(m, c) - yc(m, em) : xc(m, c) - xc(m, em);
        double dp = (h[ep] - h[c]) * (1 / dxp), dm = (h[c] - h[em]) * (1 / dxm);
        sg[dir * n + c] = .5 * (c_sgn(dp, p->tiny) + c_sgn(dm, p->tiny));
      }
    double prev[3] = {h[c], hu[c], hv[c]}, out[3];
    double g[3] = {m->gx[c], m->gy[c], m->gz[c]};
    double dg[2] = {m->dgx[c], m->dgy[c]};
    double cv[2] = {m->kx[c], m->ky[c]};
    double zf[3] = {0, 0, 0};
    src_info si;
    or_params pt;
    const double tan_bed = cell_params(&m->mats, p, c, &pt)->tan_bed;
    update_states(p, prev, out, zf, zf, zf, zf, 0., 0., dt, dc[c].d, dc[c].d + 3, cv,
                  p->sin_int, tan_bed, g, dg, dc[c].k, ft, &si);
    stop[c] = si.stop[0];
    stop[n + c] = si.stop[1];
  }
}

void or_dstep_dphi(const or_mesh* m, const or_params* p, double ft, const double* h,
                   const double* hu, const double* hv, double dt, double* dU) {
  int n = m->nx * m->ny;
  const double* U[3] = {h, hu, hv};
  double* d = (double*)ws(WS_D, sizeof(double) * 6 * n, 0);
  double* k = (double*)ws(WS_K, sizeof(double) * n, 0);
  double* dk = (double*)ws(WS_DK, sizeof(double) * n, 0);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; c++) {
    double hs[2][5];
    slopes_dual(m, p->gamma, p->tiny, U, c, d + 6 * c, hs);
    or_params t;
    const or_params* pc = cell_params(&m->mats, p, c, &t);
    k[c] = getcoef(pc, h[c], hu[c], hv[c], d + 6 * c);
    dk[c] = dgetcoef_dphi(pc, h[c], hu[c], hv[c], d + 6 * c);
  }
  double dtdx = dt / m->dx, dtdy = dt / m->dy;
#pragma omp parallel for schedule(static)
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
      or_params pt;
      const double tbc = cell_params(&m->mats, p, c, &pt)->tan_bed;
      double prev[3] = {h[c], hu[c], hv[c]}, o2[3], zf[3] = {0, 0, 0};
      double g[3] = {m->gx[c], m->gy[c], m->gz[c]}, dg[2] = {m->dgx[c], m->dgy[c]};
      double cv[2] = {m->kx[c], m->ky[c]};
      src_info si;
      update_states(p, prev, o2, zf, zf, zf, zf, dtdx, dtdy, dt, d + 6 * c, d + 6 * c + 3, cv,
                    p->sin_int, tbc, g, dg, k[c], ft, &si);
      double sec2 = 1. + tbc * tbc;
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
#pragma omp parallel for schedule(static)
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
#pragma omp parallel for schedule(static) reduction(+ : acc)
    for (int c = 0; c < n; c++)
      for (int v = 0; v < 3; v++) acc += lam_next[v * n + c] * dU[v * n + c];
    *dJdphi += acc;
  }
}
