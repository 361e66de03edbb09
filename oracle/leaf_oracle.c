/*
 * leaf_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never shipped).
 *
 * FP64 CPU restatement of the reference's order-1 forward step and its
 * per-step discrete adjoint on an adaptive mesh given as leaf tables.  It is
 * organised element by element the way the reference is: side nodes that
 * hold the face fluxes (shared between equal leaves, separate on hanging
 * sides), the Case I / II / III edge-state logic of calc_edge_states and
 * DualElem::calc_flux, a FluxJac container per element written by whichever
 * element the reference lets write it, nine Jacobian slots and the adjoint
 * gather through which_neighbor.  The product's kernels compute the same
 * quantities cell by cell from face rules, so this is an independent
 * restatement.  Expression trees follow the reference operand for operand;
 * build with -ffp-contract=off.
 */
#include "leaf_oracle.h"
#include "oracle_math.h"

#include <stdlib.h>
#include <string.h>

/* Side-node kinds (Node::info of node_key[s+4], constant.h:39-43): no
 * neighbour (domain boundary), an equal neighbour (SIDE), a coarser
 * neighbour (S_S_CON: this leaf's side lies on half of the neighbour's
 * side), two finer neighbours (S_C_CON: the side node is their corner). */
enum { K_BND = 0, K_SAME = 1, K_COARSER = 2, K_FINER = 3 };

/* positive_x_side = 1: x+ is side 1 (east), y+ side 2 (north), x- side 3,
 * y- side 0 (correct(), corrector.C:45-47; calc_edge_states zp/zm) */
static const int PLUS[2] = {1, 2}, MINUS[2] = {3, 0};

static int kind_of(const or_leaf_mesh* m, int e, int s) {
  if (m->nproc[8 * e + s] == -1) return K_BND;
  if (m->nproc[8 * e + s + 4] >= 0) return K_FINER;
  return (m->gen[m->nbr[8 * e + s]] < m->gen[e]) ? K_COARSER : K_SAME;
}

/* Element::which_neighbor, element2.C:574-583 */
static int which_neighbor(const or_leaf_mesh* m, int f, int e) {
  for (int i = 0; i < 8; i++)
    if (m->nbr[8 * f + i] == e && m->nproc[8 * f + i] >= 0) return i;
  abort();  /* inconsistent tables */
}

static int on_boundary(const or_leaf_mesh* m, int e) {
  for (int j = 0; j < 4; j++)
    if (m->nproc[8 * e + j] == -1) return 1;
  return 0;
}

/* Side nodes: node[4e + s] is the node holding the flux of side s of leaf
 * e; an equal neighbour shares it, hanging sides have their own. */
static int* build_nodes(const or_leaf_mesh* m, int* nnode) {
  int* node = (int*)malloc(sizeof(int) * 4 * (size_t)m->n);
  for (int i = 0; i < 4 * m->n; i++) node[i] = -1;
  int k = 0;
  for (int e = 0; e < m->n; e++)
    for (int s = 0; s < 4; s++) {
      if (node[4 * e + s] >= 0) continue;
      node[4 * e + s] = k;
      if (kind_of(m, e, s) == K_SAME) node[4 * m->nbr[8 * e + s] + (s + 2) % 4] = k;
      k++;
    }
  *nnode = k;
  return node;
}

/* ---- slopes ------------------------------------------------------------ */

/* Element::get_slopes, element2.C:830-939: a second neighbour on a side
 * (node S_C_CON) is averaged into that side's difference, both over the
 * centre distance to the first neighbour. */
static void slopes_fwd(const or_leaf_mesh* m, const or_params* p, const double* U[3], int e,
                       double d[6]) {
  if (on_boundary(m, e)) {
    for (int k = 0; k < 6; k++) d[k] = 0.;
    return;
  }
  for (int dir = 0; dir < 2; dir++) {
    int xp = PLUS[dir], xm = MINUS[dir];
    int ep = m->nbr[8 * e + xp], em = m->nbr[8 * e + xm];
    int ep2 = (kind_of(m, e, xp) == K_FINER) ? m->nbr[8 * e + xp + 4] : -1;
    int em2 = (kind_of(m, e, xm) == K_FINER) ? m->nbr[8 * e + xm + 4] : -1;
    double dxp = m->xc[2 * ep + dir] - m->xc[2 * e + dir];
    double dxm = m->xc[2 * e + dir] - m->xc[2 * em + dir];
    for (int j = 0; j < NV; j++) {
      double dp = (U[j][ep] - U[j][e]) / dxp;
      if (ep2 >= 0) dp = .5 * (dp + (U[j][ep2] - U[j][e]) / dxp);
      double dm = (U[j][e] - U[j][em]) / dxm;
      if (em2 >= 0) dm = .5 * (dm + (U[j][e] - U[j][em2]) / dxm);
      double dc = (dp * dxm + dm * dxp) / (dxm + dxp);
      d[j + NV * dir] = .5 * (c_sgn(dp, p->tiny) + c_sgn(dm, p->tiny)) *
                        dmin3(p->gamma * dabs(dp), p->gamma * dabs(dm), dabs(dc));
    }
  }
}

/* DualElem::get_slopes_prev, dualmesh.C:1393-1626: reciprocals, and
 * hslope_sens(dir, {self, minus1, plus1, minus2, plus2}). */
static void slopes_dual(const or_leaf_mesh* m, const or_params* p, const double* U[3], int e,
                        double d[6], double hs[2][5]) {
  if (on_boundary(m, e)) {
    for (int k = 0; k < 6; k++) d[k] = 0.;
    for (int a = 0; a < 2; a++)
      for (int k = 0; k < 5; k++) hs[a][k] = 0.;
    return;
  }
  double gamma = p->gamma;
  for (int dir = 0; dir < 2; dir++) {
    int xp = PLUS[dir], xm = MINUS[dir];
    int ep = m->nbr[8 * e + xp], em = m->nbr[8 * e + xm];
    int ep2 = (kind_of(m, e, xp) == K_FINER) ? m->nbr[8 * e + xp + 4] : -1;
    int em2 = (kind_of(m, e, xm) == K_FINER) ? m->nbr[8 * e + xm + 4] : -1;
    double dxp = m->xc[2 * ep + dir] - m->xc[2 * e + dir];
    double dxm = m->xc[2 * e + dir] - m->xc[2 * em + dir];
    double inv_dxp = 1 / dxp, inv_dxm = 1 / dxm, inv_s = 1 / (dxm + dxp);
    for (int j = 0; j < NV; j++) {
      double dp = (U[j][ep] - U[j][e]) * inv_dxp;
      if (ep2 >= 0) dp = .5 * (dp + (U[j][ep2] - U[j][e]) * inv_dxp);
      double dm = (U[j][e] - U[j][em]) * inv_dxm;
      if (em2 >= 0) dm = .5 * (dm + (U[j][e] - U[j][em2]) * inv_dxm);
      double dc = (dp * dxm + dm * dxp) * inv_s;
      double mn = dmin3(gamma * dabs(dp), gamma * dabs(dm), dabs(dc));
      d[j + NV * dir] = .5 * (c_sgn(dp, p->tiny) + c_sgn(dm, p->tiny)) * mn;
      if (j != 0) continue;
      double* s = hs[dir];
      if (.5 * (c_sgn(dp, p->tiny) + c_sgn(dm, p->tiny)) == 0) {
        for (int k = 0; k < 5; k++) s[k] = 0.;
      } else if (mn == gamma * dabs(dp)) {
        s[0] = -inv_dxp; s[1] = 0.; s[2] = inv_dxp; s[3] = 0.; s[4] = 0.;
        if (ep2 >= 0) s[2] = s[4] = .5 * inv_dxp;
      } else if (mn == gamma * dabs(dm)) {
        s[0] = inv_dxm; s[1] = -inv_dxm; s[2] = 0.; s[3] = 0.; s[4] = 0.;
        if (em2 >= 0) s[1] = s[3] = -.5 * inv_dxm;
      } else {
        s[0] = inv_dxm - inv_dxp; s[1] = -inv_dxm; s[2] = inv_dxp; s[3] = 0.; s[4] = 0.;
        if (ep2 >= 0) s[2] = s[4] = .5 * inv_dxp;
        if (em2 >= 0) s[1] = s[3] = -.5 * inv_dxm;
        s[1] *= dxp * inv_s;
        s[2] *= dxm * inv_s;
        s[3] *= dxp * inv_s;
        s[4] *= dxm * inv_s;
      }
    }
  }
}

void orl_slopes(const or_leaf_mesh* m, const or_params* p, const double* h, const double* hu,
                const double* hv, double* d) {
  const double* U[3] = {h, hu, hv};
  for (int e = 0; e < m->n; e++) slopes_fwd(m, p, U, e, d + 6 * e);
}

/* ---- forward step ------------------------------------------------------ */

static void hfv_of(const or_leaf_mesh* m, const or_params* p, const double* U[3],
                   const double* kact, int e, int dir, double f[3][3]) {
  hfv_fwd(p, dir, U[0][e], U[1][e], U[2][e], kact[e], m->gz[e], f);
}

/* calc_edge_states (element2.C:2744-3072) for one process, then the
 * outflow copies of Element::boundary_flux (element2.C:3374-3411): the flux
 * of every side node, nf[NV * node[4e + s]]. */
static void edge_states(const or_leaf_mesh* m, const or_params* p, const double* U[3],
                        const double* kact, const int* node, double* nf) {
  const int n = m->n;
  int* lst[2];
  int nl[2] = {0, 0};
  lst[0] = (int*)malloc(sizeof(int) * 2 * (size_t)n);
  lst[1] = (int*)malloc(sizeof(int) * 2 * (size_t)n);
  for (int e = 0; e < n; e++) {
    for (int side = 0; side < 2; side++) {
      int zp = PLUS[side], zm = MINUS[side];
      int kp = kind_of(m, e, zp);
      if (kp == K_BND) {
        lst[side][nl[side]++] = e;
      } else {
        double hfv[3][3], hfv1[3][3], hfv2[3][3];
        int elm1 = m->nbr[8 * e + zp];
        double* np = nf + NV * node[4 * e + zp];
        hfv_of(m, p, U, kact, e, side, hfv);
        hfv_of(m, p, U, kact, elm1, side, hfv1);
        riemann(hfv, hfv1, np);
        if (kp == K_COARSER) { /* Case II: this leaf is one of two on elm1's side */
          int zelmpos = which_neighbor(m, elm1, e);
          double* nm1 = nf + NV * node[4 * elm1 + zelmpos % 4];
          int elm2 = m->nbr[8 * elm1 + (zelmpos + 4) % 8];
          hfv_of(m, p, U, kact, elm1, side, hfv1);
          hfv_of(m, p, U, kact, elm2, side, hfv2);
          int zp2 = which_neighbor(m, elm2, elm1) % 4;
          double* np2 = nf + NV * node[4 * elm2 + zp2];
          riemann(hfv2, hfv1, np2);
          for (int v = 0; v < NV; v++) nm1[v] = 0.5 * (np[v] + np2[v]);
        } else if (kp == K_FINER) { /* Case III: two finer leaves on this side */
          int zelmpos = which_neighbor(m, elm1, e) % 4;
          double* nm1 = nf + NV * node[4 * elm1 + zelmpos];
          int elm2 = m->nbr[8 * e + zp + 4];
          hfv_of(m, p, U, kact, e, side, hfv);
          hfv_of(m, p, U, kact, elm2, side, hfv2);
          double* nm2 = nf + NV * node[4 * elm2 + which_neighbor(m, elm2, e) % 4];
          riemann(hfv, hfv2, nm2);
          for (int v = 0; v < NV; v++) {
            nm1[v] = np[v];
            np[v] = 0.5 * (nm1[v] + nm2[v]);
          }
        }
      }
      if (kind_of(m, e, zm) == K_BND) lst[side][nl[side]++] = e;
    }
  }
  /* Element::boundary_flux, element2.C:3374-3411: outflow copies */
  for (int side = 0; side < 2; side++)
    for (int q = 0; q < nl[side]; q++) {
      int e = lst[side][q], zp = PLUS[side], zm = MINUS[side];
      double* np = nf + NV * node[4 * e + zp];
      double* nm = nf + NV * node[4 * e + zm];
      if (kind_of(m, e, zp) == K_BND)
        for (int v = 0; v < NV; v++) np[v] = nm[v];
      if (kind_of(m, e, zm) == K_BND)
        for (int v = 0; v < NV; v++) nm[v] = np[v];
    }

  free(lst[0]);
  free(lst[1]);
}

int orl_step(const or_leaf_mesh* m, const or_params* p, double frict_tiny, double* h,
             double* hu, double* hv, double dt_in, double* time, double* dt_out,
             double* mass_out) {
  const int n = m->n;
  const double* U[3] = {h, hu, hv};
  double* d = (double*)malloc(sizeof(double) * 6 * (size_t)n);
  double* kact = (double*)malloc(sizeof(double) * (size_t)n);
  double* Un = (double*)malloc(sizeof(double) * 3 * (size_t)n);
  int nnode;
  int* node = build_nodes(m, &nnode);
  double* nf = (double*)calloc((size_t)nnode * NV, sizeof(double));
  int rc = 0;

  /* slopes() + get_coef_and_eigen() (get_coef_and_eigen.C:29-180) */
  double min_ev = 10000000, hmax = 0;
  for (int e = 0; e < n; e++) {
    slopes_fwd(m, p, U, e, d + 6 * e);
    or_params t;
    kact[e] = (h[e] > p->tiny)
                  ? getcoef(cell_params(&m->mats, p, e, &t), h[e], hu[e], hv[e], d + 6 * e)
                  : p->epsilon;
    if (h[e] > p->tiny) {
      if (hmax < h[e]) hmax = h[e];
      double ev = eigen(p, h[e], hu[e], hv[e], kact[e], m->gz[e]);
      min_ev = dmin2(dmin2(m->cdx[2 * e], m->cdx[2 * e + 1]) / ev, min_ev);
    }
  }
  double dt;
  if (!isnan(dt_in)) { /* a replayed dt, any value */
    dt = dt_in;
  } else {
    double dt0 = p->cfl_inner * min_ev;
    double dt1 = -0.9 * sqrt(hmax * p->epsilon * p->gscale / 9.8);
    dt = p->cfl_outer * dmin2(dt0, -dt1);
    if (dt == 0.0) dt = 0.5 * HUGE_VAL;
    dt = incrtime(p, *time, dt);
  }
  *time += dt;
  if (dt_out) *dt_out = dt;

  edge_states(m, p, U, kact, node, nf);

  /* correct() (corrector.C:36-140) -> update_states */
  double mass = 0.0;
  for (int e = 0; e < n; e++) {
    double prev[3] = {h[e], hu[e], hv[e]}, out[3];
    double g[3] = {m->gx[e], m->gy[e], m->gz[e]};
    double dg[2] = {m->dgx[e], m->dgy[e]};
    double cv[2] = {m->kx[e], m->ky[e]};
    double dtdx = dt / m->cdx[2 * e], dtdy = dt / m->cdx[2 * e + 1];
    double kc = (prev[0] > p->tiny) ? kact[e] : p->epsilon;
    src_info si;
    or_params pt;
    update_states(p, prev, out, nf + NV * node[4 * e + 1], nf + NV * node[4 * e + 2],
                  nf + NV * node[4 * e + 3], nf + NV * node[4 * e + 0], dtdx, dtdy, dt,
                  d + 6 * e, d + 6 * e + 3, cv, p->sin_int,
                  cell_params(&m->mats, p, e, &pt)->tan_bed, g, dg, kc, frict_tiny, &si);
    for (int v = 0; v < NV; v++) {
      if (!isfinite(out[v])) rc = -5;
      Un[(size_t)v * n + e] = out[v];
    }
    mass += m->cdx[2 * e] * m->cdx[2 * e + 1] * out[0];
  }
  memcpy(h, Un, sizeof(double) * n);
  memcpy(hu, Un + n, sizeof(double) * n);
  memcpy(hv, Un + 2 * (size_t)n, sizeof(double) * n);
  if (mass_out) *mass_out = mass;
  free(d); free(kact); free(Un); free(node); free(nf);
  return rc;
}

/* ---- error grid (error_grid_routines.C:13-53, error_compute.C:24-100) ---- */

/* The error grid's predicted state from the reconstructed previous state bp
 * (update_bilinear_error_grid + the update_states call of error_compute):
 * ErrorElem::get_slopes_prev (dualmesh.C:3145-3279: the dual slopes'
 * arithmetic on bp), gmfggetcoef for every element (its dry branch gives
 * epsilon), the edge states of ErrorElem::xdirflux / ydirflux
 * (dualmesh.C:3308-3387: the forward flux of bp), update_states with the
 * given dt and kact (effect_bedfrict = bedfrict[material],
 * element2.C:4008-4018). */
void orl_error_update(const or_leaf_mesh* m, const or_params* p, double frict_tiny,
                      const double* h, const double* hu, const double* hv, double dt,
                      double* out_h, double* out_hu, double* out_hv) {
  const int n = m->n;
  const double* U[3] = {h, hu, hv};
  double* d = (double*)malloc(sizeof(double) * 6 * (size_t)n);
  double* kact = (double*)calloc((size_t)n + 1, sizeof(double));
  int nnode;
  int* node = build_nodes(m, &nnode);
  double* nf = (double*)calloc((size_t)nnode * NV, sizeof(double));
  for (int e = 0; e < n; e++) {
    double hs[2][5];
    or_params t;
    slopes_dual(m, p, U, e, d + 6 * e, hs);
    kact[e] = getcoef(cell_params(&m->mats, p, e, &t), h[e], hu[e], hv[e], d + 6 * e);
  }
  edge_states(m, p, U, kact, node, nf);
  for (int e = 0; e < n; e++) {
    double prev[3] = {h[e], hu[e], hv[e]}, o[3];
    double g[3] = {m->gx[e], m->gy[e], m->gz[e]};
    double dg[2] = {m->dgx[e], m->dgy[e]};
    double cv[2] = {m->kx[e], m->ky[e]};
    double dtdx = dt / m->cdx[2 * e], dtdy = dt / m->cdx[2 * e + 1];
    src_info si;
    or_params pt;
    update_states(p, prev, o, nf + NV * node[4 * e + 1], nf + NV * node[4 * e + 2],
                  nf + NV * node[4 * e + 3], nf + NV * node[4 * e + 0], dtdx, dtdy, dt,
                  d + 6 * e, d + 6 * e + 3, cv, p->sin_int,
                  cell_params(&m->mats, p, e, &pt)->tan_bed, g, dg, kact[e], frict_tiny, &si);
    out_h[e] = o[0];
    out_hu[e] = o[1];
    out_hv[e] = o[2];
  }
  free(d); free(kact); free(node); free(nf);
}

/* ---- dual grid: FluxJac and Jacobian blocks ------------------------------ */

typedef struct leaf_dual {
  double d[6];
  double hs[2][5];
  double k;
  M3 fj[2][2][3]; /* FluxJac (side x/y)(dir minus/plus)(index self/n1/n2) */
  M3 J[9];        /* jacobianMat slots 0..8 */
} leaf_dual;

static void mset(M3 dst, const M3 src) { memcpy(dst, src, sizeof(M3)); }
/* FluxJac::set(side, dir, idx, jac1, jac2), matrix.h:228-234 */
static void mset2(M3 dst, const M3 a, const M3 b) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) dst[i][j] = (a[i][j] + b[i][j]) * .5;
}

typedef struct dhfv_t { M3 hfv, fj, sj; } dhfv_t;

static void dual_of(const or_leaf_mesh* m, const or_params* p, const double* U[3],
                    const leaf_dual* dc, int e, int dir, dhfv_t* r) {
  dual_hfv(p, dir, U[0][e], U[1][e], U[2][e], dc[e].k, m->gz[e], r->hfv, r->fj, r->sj);
}

/* update_dual_grid (dual_grid_routines.C:15-93), calc_flux
 * (DualElem::calc_flux dualmesh.C:1853-2239, boundary_flux :2241-2279) and
 * calc_jacobian (calc_jacobain.C:40-330, compute_jacobian_elem.C:15-121). */
static leaf_dual* build_dual(const or_leaf_mesh* m, const or_params* p, double ft,
                             const double* U[3], double dt) {
  const int n = m->n;
  leaf_dual* dc = (leaf_dual*)calloc((size_t)n, sizeof(leaf_dual));
  int* lst[2];
  int nl[2] = {0, 0};
  lst[0] = (int*)malloc(sizeof(int) * 2 * (size_t)n);
  lst[1] = (int*)malloc(sizeof(int) * 2 * (size_t)n);
  for (int e = 0; e < n; e++) {
    slopes_dual(m, p, U, e, dc[e].d, dc[e].hs);
    or_params t;
    dc[e].k = getcoef(cell_params(&m->mats, p, e, &t), U[0][e], U[1][e], U[2][e], dc[e].d);
  }
  for (int e = 0; e < n; e++) {
    for (int side = 0; side < 2; side++) {
      int zp = PLUS[side], zm = MINUS[side];
      int kp = kind_of(m, e, zp);
      if (kp == K_BND) {
        lst[side][nl[side]++] = e;
      } else {
        dhfv_t a, b1, b2;
        M3 jac, jac_neigh1, jac_neigh2, jac_res;
        int elm1 = m->nbr[8 * e + zp];
        dual_of(m, p, U, dc, e, side, &a);
        dual_of(m, p, U, dc, elm1, side, &b1);
        dual_riemann(a.hfv, b1.hfv, a.fj, b1.fj, a.sj, b1.sj, jac, jac_neigh1);
        mset(dc[e].fj[side][1][0], jac);
        mset(dc[e].fj[side][1][1], jac_neigh1);
        mset(dc[e].fj[side][1][2], ZERO3);
        mset(dc[elm1].fj[side][0][0], jac_neigh1);
        mset(dc[elm1].fj[side][0][1], jac);
        mset(dc[elm1].fj[side][0][2], ZERO3);
        if (kp == K_COARSER) { /* Case II */
          int zelmpos = which_neighbor(m, elm1, e);
          int elm2 = m->nbr[8 * elm1 + (zelmpos + 4) % 8];
          dual_of(m, p, U, dc, elm1, side, &b1);
          dual_of(m, p, U, dc, elm2, side, &b2);
          dual_riemann(b2.hfv, b1.hfv, b2.fj, b1.fj, b2.sj, b1.sj, jac_neigh2, jac_res);
          mset2(dc[elm1].fj[side][0][0], jac_neigh1, jac_res);
          if (which_neighbor(m, elm1, e) < 4) {
            mset2(dc[elm1].fj[side][0][1], jac, ZERO3);
            mset2(dc[elm1].fj[side][0][2], jac_neigh2, ZERO3);
          } else {
            mset2(dc[elm1].fj[side][0][1], jac_neigh2, ZERO3);
            mset2(dc[elm1].fj[side][0][2], jac, ZERO3);
          }
        } else if (kp == K_FINER) { /* Case III */
          int elm2 = m->nbr[8 * e + zp + 4];
          dual_of(m, p, U, dc, e, side, &a);
          dual_of(m, p, U, dc, elm2, side, &b2);
          dual_riemann(a.hfv, b2.hfv, a.fj, b2.fj, a.sj, b2.sj, jac_res, jac_neigh2);
          mset2(dc[e].fj[side][1][0], jac, jac_res);
          mset2(dc[e].fj[side][1][1], jac_neigh1, ZERO3);
          mset2(dc[e].fj[side][1][2], jac_neigh2, ZERO3);
          mset(dc[elm1].fj[side][0][0], jac_neigh1);
          mset(dc[elm1].fj[side][0][1], jac);
          mset(dc[elm1].fj[side][0][2], ZERO3);
          mset(dc[elm2].fj[side][0][0], jac_neigh2);
          mset(dc[elm2].fj[side][0][1], jac_res);
          mset(dc[elm2].fj[side][0][2], ZERO3);
        }
      }
      if (kind_of(m, e, zm) == K_BND) lst[side][nl[side]++] = e;
    }
  }
  for (int side = 0; side < 2; side++)
    for (int q = 0; q < nl[side]; q++) {
      int e = lst[side][q];
      if (kind_of(m, e, PLUS[side]) == K_BND)
        for (int i = 0; i < 3; i++) mset(dc[e].fj[side][1][i], dc[e].fj[side][0][i]);
      if (kind_of(m, e, MINUS[side]) == K_BND)
        for (int i = 0; i < 3; i++) mset(dc[e].fj[side][0][i], dc[e].fj[side][1][i]);
    }

  /* calc_jacobian */
  for (int e = 0; e < n; e++) {
    double prev[3] = {U[0][e], U[1][e], U[2][e]}, out[3];
    double g[3] = {m->gx[e], m->gy[e], m->gz[e]};
    double dg[2] = {m->dgx[e], m->dgy[e]};
    double cv[2] = {m->kx[e], m->ky[e]};
    double zf[3] = {0, 0, 0};
    double dtdx = dt / m->cdx[2 * e], dtdy = dt / m->cdx[2 * e + 1];
    src_info si;
    or_params pt;
    const double tan_bed = cell_params(&m->mats, p, e, &pt)->tan_bed;
    /* orgSrcSgn and stop do not depend on the fluxes */
    update_states(p, prev, out, zf, zf, zf, zf, dtdx, dtdy, dt, dc[e].d, dc[e].d + 3, cv,
                  p->sin_int, tan_bed, g, dg, dc[e].k, ft, &si);
    for (int eff = 0; eff < 9; eff++) {
      M3* J = &dc[e].J[eff];
      int zero = (eff == 0 && prev[0] == 0.);
      if (eff > 4 && m->nproc[8 * e + eff - 1] == -2) zero = 1;
      if (eff > 0) {
        int nb = m->nbr[8 * e + eff - 1];
        if (nb < 0 || U[0][nb] == 0.) zero = 1; /* void_neigh_elem */
        if (m->nproc[8 * e + eff - 1] == -1) zero = 1;
      }
      if (zero) {
        mset(*J, ZERO3);
        continue;
      }
      /* set_fluxes_hsens, calc_jacobain.C:187-330 */
      const double (*nxm)[3] = ZERO3, (*pxm)[3] = ZERO3, (*nym)[3] = ZERO3, (*pym)[3] = ZERO3;
      double dh[2] = {0., 0.};
      if (eff == 0) {
        nxm = dc[e].fj[0][0][0]; pxm = dc[e].fj[0][1][0];
        nym = dc[e].fj[1][0][0]; pym = dc[e].fj[1][1][0];
        dh[0] = dc[e].hs[0][0]; dh[1] = dc[e].hs[1][0];
      } else {
        int idx = (eff <= 4) ? 1 : 2;
        int dir = (eff - 1) % 4, opos = (dir + 2) % 4;
        int boundary = m->nproc[8 * e + opos] == -1;
        if (dir == 1) {        /* xp */
          dh[0] = dc[e].hs[0][idx == 1 ? 2 : 4];
          pxm = dc[e].fj[0][1][idx];
          if (boundary) nxm = dc[e].fj[0][0][idx];
        } else if (dir == 3) { /* xm */
          dh[0] = dc[e].hs[0][idx == 1 ? 1 : 3];
          nxm = dc[e].fj[0][0][idx];
          if (boundary) pxm = dc[e].fj[0][1][idx];
        } else if (dir == 2) { /* yp */
          dh[1] = dc[e].hs[1][idx == 1 ? 2 : 4];
          pym = dc[e].fj[1][1][idx];
          if (boundary) nym = dc[e].fj[1][0][idx];
        } else {               /* ym */
          dh[1] = dc[e].hs[1][idx == 1 ? 1 : 3];
          nym = dc[e].fj[1][0][idx];
          if (boundary) pym = dc[e].fj[1][1][idx];
        }
      }
      /* calc_jacobian_elem, compute_jacobian_elem.C:15-121 */
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          (*J)[i][j] = +dtdx * (pxm[i][j] - nxm[i][j]) + dtdy * (pym[i][j] - nym[i][j]);
      if (eff == 0)
        for (int i = 0; i < 3; i++) (*J)[i][i] -= 1.;
      if (prev[0] > p->tiny) {
        double kact = dc[e].k;
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
                              - alpha * g[2] * dc[e].d[3]);
          if (si.sgn[2] != 0. && !si.stop[0]) {
            (*J)[1][0] -= -dt * si.sgn[2] * tb * (g[2] - vx_sq * cv[0]);
            (*J)[1][1] -= -dt * si.sgn[2] * tb * (2 * vx * cv[0]);
            (*J)[1][2] -= 0.;
          }
          (*J)[2][0] -= dt * (g[1] - betta * prev[0] * (2 * dg[0] + g[2] * dh[0])
                              - betta * g[2] * dc[e].d[0]);
          if (si.sgn[3] != 0. && !si.stop[1]) {
            (*J)[2][0] -= -dt * si.sgn[3] * tb * (g[2] - vy_sq * cv[1]);
            (*J)[2][1] -= 0.;
            (*J)[2][2] -= -dt * si.sgn[3] * tb * (2 * vy * cv[1]);
          }
        }
      }
    }
  }
  free(lst[0]); free(lst[1]);
  return dc;
}

void orl_jacobian(const or_leaf_mesh* m, const or_params* p, double frict_tiny,
                  const double* h, const double* hu, const double* hv, double dt, double* Jout) {
  const double* U[3] = {h, hu, hv};
  leaf_dual* dc = build_dual(m, p, frict_tiny, U, dt);
  for (int e = 0; e < m->n; e++)
    for (int s = 0; s < 9; s++)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Jout[(size_t)e * 81 + s * 9 + i * 3 + j] = dc[e].J[s][i][j];
  free(dc);
}

/* ---- d U / d phi_bed at frozen dt (no reference counterpart) ------------ */

/* Flux of side s of leaf e and its phi_bed derivative, by the forward rules:
 * HLL with the minus-side leaf on the left; a side with two finer leaves
 * takes half the sum of the two small-face fluxes; a boundary side copies
 * the opposite side (outflow). */
static void side_flux(const or_leaf_mesh* m, const or_params* p, const double* U[3],
                      const double* k, const double* dk, int e, int s, double F[3],
                      double dF[3]) {
  int kd = kind_of(m, e, s);
  if (kd == K_BND) {
    side_flux(m, p, U, k, dk, e, (s + 2) % 4, F, dF);
    return;
  }
  int dir = (s == 1 || s == 3) ? 0 : 1, plus = (s == 1 || s == 2);
  int cnt = (kd == K_FINER) ? 2 : 1;
  double Fi[2][3], dFi[2][3];
  for (int q = 0; q < cnt; q++) {
    int f = m->nbr[8 * e + s + 4 * q];
    int l = plus ? e : f, r = plus ? f : e;
    double hl[3][3], hr[3][3], dFl[3], dWl[3], dFr[3], dWr[3];
    hfv_fwd(p, dir, U[0][l], U[1][l], U[2][l], k[l], m->gz[l], hl);
    hfv_fwd(p, dir, U[0][r], U[1][r], U[2][r], k[r], m->gz[r], hr);
    riemann(hl, hr, Fi[q]);
    dhfv_dk(p, dir, U[0][l], k[l], m->gz[l], dFl, dWl);
    dhfv_dk(p, dir, U[0][r], k[r], m->gz[r], dFr, dWr);
    for (int v = 0; v < 3; v++) {
      dFl[v] *= dk[l]; dWl[v] *= dk[l]; dFr[v] *= dk[r]; dWr[v] *= dk[r];
    }
    driemann(hl, hr, dFl, dWl, dFr, dWr, dFi[q]);
  }
  for (int v = 0; v < 3; v++) {
    F[v] = (cnt == 2) ? 0.5 * (Fi[0][v] + Fi[1][v]) : Fi[0][v];
    dF[v] = (cnt == 2) ? 0.5 * (dFi[0][v] + dFi[1][v]) : dFi[0][v];
  }
}

void orl_dstep_dphi(const or_leaf_mesh* m, const or_params* p, double ft, const double* h,
                    const double* hu, const double* hv, double dt, double* dU) {
  const int n = m->n;
  const double* U[3] = {h, hu, hv};
  double* d = (double*)malloc(sizeof(double) * 6 * (size_t)n);
  double* k = (double*)malloc(sizeof(double) * (size_t)n);
  double* dk = (double*)malloc(sizeof(double) * (size_t)n);
  double hs[2][5];
  for (int e = 0; e < n; e++) {
    slopes_dual(m, p, U, e, d + 6 * e, hs);
    or_params t;
    const or_params* pc = cell_params(&m->mats, p, e, &t);
    k[e] = getcoef(pc, h[e], hu[e], hv[e], d + 6 * e);
    dk[e] = dgetcoef_dphi(pc, h[e], hu[e], hv[e], d + 6 * e);
  }
  for (int e = 0; e < n; e++) {
    double flux[4][3], dflux[4][3]; /* E, N, W, S */
    static const int side_of[4] = {1, 2, 3, 0};
    for (int f = 0; f < 4; f++) side_flux(m, p, U, k, dk, e, side_of[f], flux[f], dflux[f]);
    double dtdx = dt / m->cdx[2 * e], dtdy = dt / m->cdx[2 * e + 1];
    double out[3];
    for (int v = 0; v < 3; v++)
      out[v] = -dtdx * (dflux[0][v] - dflux[2][v]) - dtdy * (dflux[1][v] - dflux[3][v]);
    if (h[e] - dtdx * (flux[0][0] - flux[2][0]) - dtdy * (flux[1][0] - flux[3][0]) < 0.)
      out[0] = 0.;
    if (h[e] > p->tiny) {
      double prev[3] = {h[e], hu[e], hv[e]}, o2[3], zf[3] = {0, 0, 0};
      double g[3] = {m->gx[e], m->gy[e], m->gz[e]}, dg[2] = {m->dgx[e], m->dgy[e]};
      double cv[2] = {m->kx[e], m->ky[e]};
      src_info si;
      or_params pt;
      const double tbc = cell_params(&m->mats, p, e, &pt)->tan_bed;
      update_states(p, prev, o2, zf, zf, zf, zf, dtdx, dtdy, dt, d + 6 * e, d + 6 * e + 3, cv,
                    p->sin_int, tbc, g, dg, k[e], ft, &si);
      double sec2 = 1. + tbc * tbc;
      double vx = prev[1] / prev[0], vy = prev[2] / prev[0];
      const double* dUx = d + 6 * e;
      const double* dUy = d + 6 * e + 3;
      double ds2x = si.sgn[0] * prev[0] * dk[e] * (g[2] * dUy[0] + dg[1] * prev[0]) * p->sin_int;
      double ds3x = si.sgn[2] * umax(g[2] * prev[0] + vx * prev[1] * cv[0], 0.0) * sec2;
      double ds2y = si.sgn[1] * prev[0] * dk[e] * (g[2] * dUx[0] + dg[0] * prev[0]) * p->sin_int;
      double ds3y = si.sgn[3] * umax(g[2] * prev[0] + vy * prev[2] * cv[1], 0.0) * sec2;
      out[1] += dt * (-ds2x - ds3x);
      out[2] += dt * (-ds2y - ds3y);
    }
    dU[e] = out[0];
    dU[n + e] = out[1];
    dU[2 * (size_t)n + e] = out[2];
  }
  free(d); free(k); free(dk);
}

/* ---- adjoint step (calc_adjoint.C:35-167) ------------------------------ */

void orl_adjoint_step(const or_leaf_mesh* m, const or_params* p, double frict_tiny,
                      const double* h, const double* hu, const double* hv, double dt,
                      const double* lam_next, const double* g, double* lam_out,
                      double* dJdphi) {
  const int n = m->n;
  const double* U[3] = {h, hu, hv};
  leaf_dual* dc = build_dual(m, p, frict_tiny, U, dt);
  for (int e = 0; e < n; e++) {
    double adj[3] = {0., 0., 0.};
    for (int eff = 0; eff < 9; eff++) {
      int src = e, slot = 0;
      if (eff > 0) {
        if (eff > 4 && !(m->nproc[8 * e + eff - 1] > -2)) continue;
        src = m->nbr[8 * e + eff - 1];
        if (src < 0) continue; /* no element behind a boundary side */
        slot = which_neighbor(m, src, e) + 1;
      }
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) adj[k] += lam_next[(size_t)l * n + src] * dc[src].J[slot][l][k];
    }
    for (int j = 0; j < 3; j++) {
      double gj = g ? g[(size_t)j * n + e] : 0.;
      double v = gj - adj[j];
      if (fabs(v) < 1e-16) v = 0.;
      lam_out[(size_t)j * n + e] = v;
    }
  }
  free(dc);
  if (dJdphi) {
    double* dU = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    orl_dstep_dphi(m, p, frict_tiny, h, hu, hv, dt, dU);
    double acc = 0.;
    for (int e = 0; e < n; e++)
      for (int v = 0; v < 3; v++) acc += lam_next[(size_t)v * n + e] * dU[(size_t)v * n + e];
    *dJdphi += acc;
  }
}
