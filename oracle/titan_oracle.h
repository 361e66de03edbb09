/*
 * titan_oracle.h -- CPU restatement (FP64, plain C) of the reference's
 * order-1 forward step and per-step discrete adjoint on a uniform grid.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this
 * code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg use it, and only as the checker.
 *
 * Cell layout: c = j*nx + i (x fastest).  Directions 0=E(x+), 1=N(y+),
 * 2=W(x-), 3=S(y-); positive_x_side is taken as 0 so the reference's
 * side numbering (element2.C:694-716) maps 1:1 onto these directions.
 */
#ifndef TITAN_ORACLE_H
#define TITAN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Per-cell materials (Element::material, MatProps::bedfrict[material]):
 * cell c uses entry mat[c] of tan_bed, kact_root and dkroot (d kact_root /
 * d phi_bed); mat NULL: every cell uses or_params' values. */
typedef struct or_materials {
  const unsigned char* mat;
  const double *tan_bed, *kact_root, *dkroot;
} or_materials;

typedef struct or_mesh {
  int nx, ny;
  double dx, dy;     /* cell sizes; centre coordinates x0+i*dx, y0+j*dy */
  double x0, y0;
  /* static per-cell coefficients, each an array of nx*ny doubles */
  const double *gx, *gy, *gz;   /* gravity[0..2]            */
  const double *dgx, *dgy;      /* d_gravity[0..1]          */
  const double *kx, *ky;        /* curvature[0..1]          */
  or_materials mats;            /* per-cell bed friction (mat NULL: one)   */
} or_mesh;

typedef struct or_params {
  double sin_int;    /* sin(phi_int)                                   */
  double tan_bed;    /* tan(phi_bed)                                   */
  double kact_c1;    /* 2/cos(phi_int)^2          (getcoef.f:40)       */
  double kact_root;  /* sqrt|1-(1+tan^2 bed)cos^2 int| (getcoef.f:41)  */
  double dkroot_dphi;/* d kact_root / d phi_bed (0 if not needed)      */
  double epsilon;    /* H/L scale ratio                                */
  double gscale;     /* GRAVITY_SCALE                                  */
  double tiny;       /* GEOFLOW_TINY                                   */
  double gamma;      /* slope-limiter gamma                            */
  double cfl_inner;  /* 0.5, get_coef_and_eigen.C:168                  */
  double cfl_outer;  /* 0.5, get_coef_and_eigen.C:175                  */
  double maxtime;    /* incrtime clamp (properties.h:397-410)          */
  double nextoutput;
  double nextsave;
} or_params;

/* One forward step (step.C:29-275, order 1).  If dt_in is not NaN the step
 * uses it (replay with a recorded dt, any value); otherwise the CFL dt is computed
 * (get_coef_and_eigen.C + incrtime) and *time is advanced.  U is updated in
 * place.  Returns 0, or -5 if a non-finite state appears. */
int or_step(const or_mesh* m, const or_params* p, double frict_tiny,
            double* h, double* hu, double* hv,
            double dt_in, double* time, double* dt_out, double* mass_out);

/* Forward slopes (dU [6n]: dh/dx, dhu/dx, dhv/dx, dh/dy, dhu/dy, dhv/dy per
 * cell) and kact [n] of every cell, as or_step computes them (slopes.C:25-83,
 * element2.C:830-939, getcoef.f:19-66; kact = epsilon where h <= tiny). */
void or_kact_slopes(const or_mesh* m, const or_params* p, const double* h,
                    const double* hu, const double* hv, double* kact, double* dU);

/* CFL dt of the current state (get_coef_and_eigen.C:29-180 + incrtime). */
double or_compute_dt(const or_mesh* m, const or_params* p,
                     const double* h, const double* hu, const double* hv,
                     double time);

/* Per-cell Jacobian blocks of the residual R^{s+1} = U^{s+1} - F(U^s)
 * w.r.t. U^s (calc_jacobain.C:40-153): J[c*9*9 + slot*9 + r*3 + col],
 * slot 0 = self, 1..4 = neighbour in direction 0..3, 5..8 = second
 * neighbour (always zero on a uniform grid). */
void or_jacobian(const or_mesh* m, const or_params* p, double frict_tiny,
                 const double* h, const double* hu, const double* hv,
                 double dt, double* J);

/* One reverse step (dual_solver.C:183-248 with calc_jacobian and
 * calc_adjoint every step): given U^s and lambda^{s+1} (3 SoA arrays),
 * writes lambda^s = g^s - sum_j J_j^T lambda^{s+1}_j.  g may be NULL (zero).
 * If dJdphi is non-NULL, adds lambda^{s+1} . dF/dphi_bed (dt frozen). */
void or_adjoint_step(const or_mesh* m, const or_params* p, double frict_tiny,
                     const double* h, const double* hu, const double* hv,
                     double dt, const double* lam_next, const double* g,
                     double* lam_out, double* dJdphi);

/* Per cell: the h-slope limiter's sign factor in x and y (sg[2n]) and the
 * Coulomb stop flags (stop[2n]) of U at step s -- where the reference's
 * adjoint approximations act (test classification only). */
void or_switches(const or_mesh* m, const or_params* p, double frict_tiny,
                 const double* h, const double* hu, const double* hv, double dt,
                 double* sg, int* stop);

/* d U^{s+1} / d phi_bed for every cell at frozen dt (3 SoA arrays). */
void or_dstep_dphi(const or_mesh* m, const or_params* p, double frict_tiny,
                   const double* h, const double* hu, const double* hv,
                   double dt, double* dU);

#ifdef __cplusplus
}
#endif
#endif
