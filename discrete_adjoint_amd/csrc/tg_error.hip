// FP64 kernels of the error indicator on the error grid (the reference's
// Error build: dual_solver.C:143-175, 208-240).  The error grid is the mesh
// with every leaf refined once; its elements are given as leaf tables
// (LeafDev) plus, per element, its father leaf and the interpolation rule
// and points tg_host_error_stencil picked.  One thread per error element:
//   k_err_recon    -- bilinear_interp_elem (bilinear_interp.C:427-556): the
//                     reconstructed state, previous state and adjoint from
//                     the father leaves' values at the stencil points;
//   k_err_kact     -- update_bilinear_error_grid (error_grid_routines.C:13-
//                     53): ErrorElem::get_slopes_prev on the reconstructed
//                     previous state (the dual slopes' arithmetic,
//                     dualmesh.C:3145-3279) and gmfggetcoef for every
//                     element;
//   k_err_residual -- the edge fluxes of ErrorElem::xdirflux / ydirflux
//                     (dualmesh.C:3308-3387, calc_edge_states<ErrorElem>),
//                     update_states from the reconstructed previous state,
//                     and error_compute (error_compute.C:40-95): residual =
//                     predicted - reconstructed state, error = sum |res_j
//                     (lambda_j - lambda~_j)|, correction = sum res_j
//                     lambda_j, lambda the father's adjoint.
#include <hip/hip_runtime.h>

#include "tg_kernels.h"
#include "tg_leafops.h"
#include "tg_math.h"

namespace tg {

namespace {

__device__ __forceinline__ double recon_one(int kind, const double* px, const double* py,
                                            const double* f, double x, double y,
                                            const double* A, const int* piv, bool singular) {
  switch (kind) {
    case ERR_COPY:
      return f[0];
    case ERR_LINEAR_X:  // linear_interp (bilinear_interp.C:346-349), as written
      return (f[1] - f[0]) / (px[1] - px[0]) * (x - px[0]);
    case ERR_LINEAR_Y:
      return (f[1] - f[0]) / (py[1] - py[0]) * (y - py[0]);
    case ERR_BARYCENTRIC: {  // barycentric_interpolation (:398-408)
      const double inv_det =
          1 / ((py[1] - py[2]) * (px[0] - px[2]) + (px[2] - px[1]) * (py[0] - py[2]));
      const double l1 = ((py[1] - py[2]) * (x - px[2]) + (px[2] - px[1]) * (y - py[2])) * inv_det;
      const double l2 = ((py[2] - py[0]) * (x - px[2]) + (px[0] - px[2]) * (y - py[2])) * inv_det;
      const double l3 = 1 - l1 - l2;
      return l1 * f[0] + l2 * f[1] + l3 * f[2];
    }
    default: {  // bilinear_interpolation (:350-383): dgesv of the 4x4 fit
      if (f[0] == 0. && f[1] == 0. && f[2] == 0. && f[3] == 0.) return 0.;
      double b[4] = {f[0], f[1], f[2], f[3]};
      if (!singular) {
        // dgetrs: row interchanges, unit lower, then upper triangular solve
        // (the reference BLAS dtrsm column order)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int p = piv[i];
          const double t = b[i];
          b[i] = b[p];
          b[p] = t;
        }
#pragma unroll
        for (int k = 0; k < 4; k++)
          if (b[k] != 0.)
#pragma unroll
            for (int i = k + 1; i < 4; i++) b[i] = b[i] - b[k] * A[4 * i + k];
#pragma unroll
        for (int k = 3; k >= 0; k--)
          if (b[k] != 0.) {
            b[k] = b[k] / A[4 * k + k];
#pragma unroll
            for (int i = 0; i < k; i++) b[i] = b[i] - b[k] * A[4 * i + k];
          }
      }
      return b[0] + b[1] * x + b[2] * y + b[3] * x * y;
    }
  }
}

// LU factorisation with partial pivoting of the 4x4 matrix rows [1, x, y,
// xy] (dgetrf: the pivot is the first largest |entry|, the column below it
// is scaled by its reciprocal, the rest updated by rank-one steps).  Returns
// whether a pivot is exactly zero (dgesv then leaves the right side as is).
__device__ __forceinline__ bool lu4(const double* px, const double* py, double* A, int* piv) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    A[4 * i] = 1;
    A[4 * i + 1] = px[i];
    A[4 * i + 2] = py[i];
    A[4 * i + 3] = px[i] * py[i];
  }
  bool singular = false;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    int p = k;
    double m = fabs(A[4 * k + k]);
#pragma unroll
    for (int i = k + 1; i < 4; i++)
      if (fabs(A[4 * i + k]) > m) {
        m = fabs(A[4 * i + k]);
        p = i;
      }
    piv[k] = p;
    if (A[4 * p + k] != 0.) {
      if (p != k)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const double t = A[4 * k + j];
          A[4 * k + j] = A[4 * p + j];
          A[4 * p + j] = t;
        }
      const double r = 1. / A[4 * k + k];
#pragma unroll
      for (int i = k + 1; i < 4; i++) A[4 * i + k] = A[4 * i + k] * r;
    } else {
      singular = true;
    }
#pragma unroll
    for (int j = k + 1; j < 4; j++) {
      const double t = -A[4 * k + j];
      if (t != 0.)
#pragma unroll
        for (int i = k + 1; i < 4; i++) A[4 * i + j] = A[4 * i + j] + A[4 * i + k] * t;
    }
  }
  return singular;
}

__global__ __launch_bounds__(BLOCK) void k_err_recon(ErrArgs a) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= a.n) return;
  const int kind = a.kind[e];
  double px[4], py[4];
  int lf[4];
  const int np = kind == ERR_COPY ? 1 : (kind == ERR_BARYCENTRIC ? 3 : (kind == ERR_BILINEAR ? 4 : 2));
#pragma unroll
  for (int k = 0; k < 4; k++) {
    px[k] = a.pxy[8 * (size_t)e + 2 * k];
    py[k] = a.pxy[8 * (size_t)e + 2 * k + 1];
    lf[k] = k < np ? a.pts[4 * (size_t)e + k] : -1;
  }
  if (kind == ERR_COPY) lf[0] = a.father[e];  // the element's own (its father's) values
  const double x = a.L.xc[e], y = a.L.xc[a.n + e];
  double A[16];
  int piv[4] = {0, 1, 2, 3};
  const bool singular = kind == ERR_BILINEAR ? lu4(px, py, A, piv) : false;
  // fields: previous state, state, adjoint (3 each) of the father leaves
#pragma unroll
  for (int v = 0; v < 9; v++) {
    const double* src = a.fwd + (size_t)v * a.nfwd;
    double f[4];
#pragma unroll
    for (int k = 0; k < 4; k++) f[k] = lf[k] >= 0 ? src[lf[k]] : 0.;
    a.rec[(size_t)v * a.n + e] = recon_one(kind, px, py, f, x, y, A, piv, singular);
  }
}

// ErrorElem::get_slopes_prev of one direction on the reconstructed previous
// state: the dual slopes' arithmetic without the sensitivities
__device__ __forceinline__ double err_slope(const ErrArgs& a, const double* u, int e,
                                            const DirNb& q) {
  return lslope_dual(u, e, q, a.gamma, a.tiny, false).d;
}

__global__ __launch_bounds__(BLOCK) void k_err_kact(ErrArgs a) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= a.n) return;
  const double* bh = a.rec;
  const double* bhu = a.rec + a.n;
  const double* bhv = a.rec + 2 * (size_t)a.n;
  const uint32_t meta = a.L.meta[e];
  double hx = 0., hux = 0., hy = 0., hvy = 0.;
  if (!touches_boundary(meta)) {
    const DirNb qx = dir_nb(a.L, e, meta, 0), qy = dir_nb(a.L, e, meta, 1);
    hx = err_slope(a, bh, e, qx);
    hux = err_slope(a, bhu, e, qx);
    hy = err_slope(a, bh, e, qy);
    hvy = err_slope(a, bhv, e, qy);
  }
  const double kroot = mat_of(a.I, a.mat, a.mtab, e).kroot;
  a.kact[e] = getcoef(bh[e], bhu[e], bhv[e], hx, hux, hy, hvy, a.I.kc1, kroot, a.I.eps, a.tiny);
}

__global__ __launch_bounds__(BLOCK) void k_err_residual(ErrArgs a) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= a.n) return;
  const size_t n = a.n;
  const Fields B{a.rec, a.rec + n, a.rec + 2 * n};
  const uint32_t meta = a.L.meta[e];
  double hx = 0., hvx = 0., hy = 0., huy = 0.;
  if (!touches_boundary(meta)) {
    const DirNb qx = dir_nb(a.L, e, meta, 0), qy = dir_nb(a.L, e, meta, 1);
    hx = err_slope(a, B.h, e, qx);
    hvx = err_slope(a, B.hv, e, qx);
    hy = err_slope(a, B.h, e, qy);
    huy = err_slope(a, B.hu, e, qy);
  }
  const double h = B.h[e], hu = B.hu[e], hv = B.hv[e];
  const double k = a.kact[e], g2 = a.S.gz[e];
  const Hfv mx = hfv_fwd(0, h, hu, hv, k, g2, a.tiny);
  const Hfv my = hfv_fwd(1, h, hu, hv, k, g2, a.tiny);
  double F[4][3];
#pragma unroll
  for (int s = 0; s < 4; s++)
    if (kind_of(meta, s) != LEAF_BOUNDARY)
      side_flux(a.L, B, a.kact, a.S.gz, a.tiny, e, meta, s, (s & 1) ? mx : my, F[s]);
#pragma unroll
  for (int s = 0; s < 4; s++)  // outflow copies (Element::boundary_flux)
    if (kind_of(meta, s) == LEAF_BOUNDARY)
#pragma unroll
      for (int v = 0; v < 3; v++) F[s][v] = F[(s + 2) & 3][v];
  const double dtdx = a.dt / a.L.cdx[e], dtdy = a.dt / a.L.cdx[n + e];
  const double prev[3] = {h, hu, hv};
  double out[3];
  SrcInfo si;
  update_states(prev, out, F[1], F[2], F[3], F[0], dtdx, dtdy, a.dt, hx, hvx, hy, huy,
                a.S.kx[e], a.S.ky[e], a.I.sin_int, mat_of(a.I, a.mat, a.mtab, e).tan_bed,
                a.S.gx[e], a.S.gy[e], g2, a.S.dgx[e], a.S.dgy[e], k, a.ft, a.tiny, &si);
  const int f = a.father[e];
  double err = 0., corr = 0.;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const double res = out[j] - a.rec[(3 + j) * n + e];
    const double lam = a.fwd[(size_t)(6 + j) * a.nfwd + f];
    const double blam = a.rec[(6 + j) * n + e];
    a.res[j * n + e] = res;
    err += fabs(res * (lam - blam));
    corr += res * lam;
  }
  a.err[e] = err;
  a.corr[e] = corr;
}

}  // namespace

void launch_error(const ErrArgs& a, hipStream_t s, hipEvent_t* ev) {
  const int g = (a.n + BLOCK - 1) / BLOCK;
  if (g < 1) return;
  if (ev) hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_err_recon, dim3(g), dim3(BLOCK), 0, s, a);
  if (ev) hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(k_err_kact, dim3(g), dim3(BLOCK), 0, s, a);
  if (ev) hipEventRecord(ev[2], s);
  hipLaunchKernelGGL(k_err_residual, dim3(g), dim3(BLOCK), 0, s, a);
  if (ev) hipEventRecord(ev[3], s);
}

}  // namespace tg
