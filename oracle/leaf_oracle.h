/*
 * leaf_oracle.h -- CPU restatement (FP64, plain C) of the reference's order-1
 * forward step and per-step discrete adjoint on an adaptive (2:1 balanced)
 * mesh given as flattened leaf tables.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this
 * code; only tests/ use it, as the checker.
 *
 * Sides follow the reference's neighbour numbering on generated grids
 * (positive_x_side = 1): 0 south, 1 east, 2 north, 3 west.  Slot s holds the
 * neighbour on side s; with two finer neighbours on a side slot s holds the
 * one next to corner s and slot s+4 the one next to corner s+1.
 */
#ifndef LEAF_ORACLE_H
#define LEAF_ORACLE_H

#include "titan_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_leaf_mesh {
  int n;
  const int* nbr;     /* [8n] Element::neighbor as leaf indices (-1: boundary)   */
  const int* nproc;   /* [8n] Element::neigh_proc: -1 boundary, -2 no second
                         neighbour on that side, 0 present                      */
  const int* gen;     /* [n]  Element::generation                               */
  const double* cdx;  /* [2n] Element::dx                                       */
  const double* xc;   /* [2n] Element::coord                                    */
  const double *gx, *gy, *gz, *dgx, *dgy, *kx, *ky;  /* static coefficients [n] */
  or_materials mats;  /* per-leaf bed friction (mat NULL: one material)       */
} or_leaf_mesh;

/* One forward step (step.C:29-275, order 1), U updated in place.  dt_in not
 * NaN replays that dt (any value); otherwise the CFL dt is computed and *time advanced.
 * Returns 0, -5 on a non-finite state. */
int orl_step(const or_leaf_mesh* m, const or_params* p, double frict_tiny, double* h,
             double* hu, double* hv, double dt_in, double* time, double* dt_out,
             double* mass_out);

/* Jacobian blocks J[c*81 + slot*9 + r*3 + col] of R^{s+1} w.r.t. U^s
 * (calc_jacobain.C:40-153): slot 0 self, slot k in 1..8 the neighbour in
 * Element::neighbor[k-1]. */
void orl_jacobian(const or_leaf_mesh* m, const or_params* p, double frict_tiny,
                  const double* h, const double* hu, const double* hv, double dt, double* J);

/* One reverse step: lambda^s = g - sum J^T lambda^{s+1} (calc_adjoint.C:35-167);
 * with dJdphi, adds lambda^{s+1} . dU^{s+1}/dphi_bed (frozen dt). */
void orl_adjoint_step(const or_leaf_mesh* m, const or_params* p, double frict_tiny,
                      const double* h, const double* hu, const double* hv, double dt,
                      const double* lam_next, const double* g, double* lam_out,
                      double* dJdphi);

/* dU^{s+1}/dphi_bed per cell at frozen dt (3 SoA arrays). */
void orl_dstep_dphi(const or_leaf_mesh* m, const or_params* p, double frict_tiny,
                    const double* h, const double* hu, const double* hv, double dt,
                    double* dU);

/* Error grid: the state update_states predicts from the reconstructed
 * previous state (h, hu, hv) with the error elements' slopes
 * (ErrorElem::get_slopes_prev), kact and fluxes, for the given dt
 * (error_grid_routines.C:13-53, error_compute.C:40-84). */
void orl_error_update(const or_leaf_mesh* m, const or_params* p, double frict_tiny,
                      const double* h, const double* hu, const double* hv, double dt,
                      double* out_h, double* out_hu, double* out_hv);

/* Forward limited slopes d[6n] (Element::get_slopes, element2.C:830-939). */
void orl_slopes(const or_leaf_mesh* m, const or_params* p, const double* h, const double* hu,
                const double* hv, double* d);

#ifdef __cplusplus
}
#endif
#endif
