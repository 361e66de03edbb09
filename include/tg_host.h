/*
 * tg_host.h -- host-side set-up that feeds tg_set_mesh / tg_set_state:
 * synthetic DEMs, the per-cell static coefficients the reference derives
 * from a GIS raster, and the initial pile heights.  Plain C ABI, host only.
 */
#ifndef TG_HOST_H
#define TG_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A GRASS-style elevation raster: row 0 is the northern edge (ymax), column 0
 * the western edge (xmin); values are float32 like the reference's fcell
 * reader (GisApi.C:153-285). */
typedef struct tg_dem {
  int nrows, ncols;
  double xmin, ymax, res;
  const float* elev;    /* [nrows*ncols], row-major */
} tg_dem;

/* Analytic DEM: a plane of slope angle slope_deg falling towards +x that
 * ends at x_toe and blends into a flat runout with a C1 cosine transition of
 * total width blend (no reference counterpart; synthetic input). */
int tg_host_dem_plane_runout(int nrows, int ncols, double xmin, double ymax, double res,
                             double slope_deg, double x_toe, double blend, float* elev);

/* Seeded diamond-square fractal DEM on a (2^k+1)^2 grid, rescaled so that
 * max - min == relief; roughness H (amplitude falls by 2^-H per level). */
int tg_host_dem_diamond_square(int n, double relief, double H, uint64_t seed, float* elev);

/* Per-cell static coefficients of a uniform block of cells from the DEM:
 * Horn 3x3 slope grids (calculate_slope, GisApi.C:1295-1345), bilinear slope
 * sampling (Get_slope, GisApi.C:731-762), curvature (Get_curvature,
 * GisApi.C:764-833), gravity vector (calc_gravity_vector, element2.C:3545-3564)
 * and d gravity_z (calc_d_gravity, element2.C:3574-3653).
 * grav [3*n], dgrav [2*n], curv [2*n] FP64 SoA; n = nx*ny. */
int tg_host_static_coefs(const tg_dem* dem, int nx, int ny, double dx, double dy,
                         double x0, double y0, double length_scale, double gravity_scale,
                         double* grav, double* dgrav, double* curv);

/* Elliptical paraboloid piles (elliptical_pile_height, init_piles.C:167-277,
 * PARABALOID branch): each cell takes the 9-node weighted average
 * (corners + 2*edges + 4*centre)/16 of max-over-piles heights. Rotation in
 * degrees (cos/sin of rot evaluated on the host).  hu, hv are set to zero
 * (piles start at rest). */
int tg_host_piles(int nx, int ny, double dx, double dy, double x0, double y0,
                  int npiles, const double* xcen, const double* ycen,
                  const double* major, const double* minor, const double* rot_deg,
                  const double* height, double* h, double* hu, double* hv);

/* Terminal functional weights for "h integrated over a box": w_h = dx*dy for
 * cells whose centre lies strictly inside [xmin,xmax]x[ymin,ymax]
 * (box test as MaxH_Functional::clac_func_sens, calc_adjoint.C:467-491),
 * w_hu = w_hv = 0.  w is FP64 SoA [3*n]. */
int tg_host_box_weights(int nx, int ny, double dx, double dy, double x0, double y0,
                        double xmin, double xmax, double ymin, double ymax, double* w);

#ifdef __cplusplus
}
#endif
#endif
