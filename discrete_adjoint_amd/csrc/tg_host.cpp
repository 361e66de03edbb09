// Host-side set-up: synthetic DEMs, GIS-derived static coefficients, piles.
// These run once per mesh on the CPU (the reference does them per element in
// setup_geoflow / init_piles); the results are the arrays tg_set_mesh and
// tg_set_state copy to the device.
#include "../../include/tg_host.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr double kTiny = 1.e-04;  // GEOFLOW_TINY, constant.h:68

struct Raster {
  int nrows, ncols;
  double xmin, ymax, res;
  std::vector<float> v;
  float at(int r, int c) const { return v[(size_t)r * ncols + c]; }
  float& at(int r, int c) { return v[(size_t)r * ncols + c]; }
};

// interpolate_bilinear_at, GisApi.C:648-662 (float grid, double weights)
double bilinear(double x, double y, const Raster& g) {
  int col = (int)x, row = (int)y;
  double dx = x - (double)col, dy = y - (double)row;
  double dx1 = 1.0 - dx, dy1 = 1.0 - dy;
  double p1 = g.at(row, col) * dy1 + g.at(row + 1, col) * dy;
  double p2 = g.at(row, col + 1) * dy1 + g.at(row + 1, col + 1) * dy;
  return p1 * dx1 + p2 * dx;
}

// calculate_slope, GisApi.C:1295-1345: Horn 3x3 operator in float
void horn_slopes(const Raster& e, Raster& xs, Raster& ys) {
  xs = e; ys = e;
  const int nr = e.nrows, nc = e.ncols;
  const float res8 = 8 * (float)e.res;
  for (int r = 1; r < nr - 1; r++)
    for (int c = 1; c < nc - 1; c++) {
      xs.at(r, c) = ((e.at(r - 1, c + 1) - e.at(r - 1, c - 1)) +
                     2 * (e.at(r, c + 1) - e.at(r, c - 1)) +
                     (e.at(r + 1, c + 1) - e.at(r + 1, c - 1))) / res8;
      ys.at(r, c) = ((e.at(r - 1, c - 1) - e.at(r + 1, c - 1)) +
                     2 * (e.at(r - 1, c) - e.at(r + 1, c)) +
                     (e.at(r - 1, c + 1) - e.at(r + 1, c + 1))) / res8;
    }
  for (int c = 1; c < nc - 1; c++) {
    xs.at(0, c) = xs.at(1, c); xs.at(nr - 1, c) = xs.at(nr - 2, c);
    ys.at(0, c) = ys.at(1, c); ys.at(nr - 1, c) = ys.at(nr - 2, c);
  }
  for (int r = 0; r < nr; r++) {
    xs.at(r, 0) = xs.at(r, 1); xs.at(r, nc - 1) = xs.at(r, nc - 2);
    ys.at(r, 0) = ys.at(r, 1); ys.at(r, nc - 1) = ys.at(r, nc - 2);
  }
}

// Get_slope, GisApi.C:731-762
void get_slope(const Raster& xs, const Raster& ys, double x, double y, double* zx, double* zy) {
  x = (x - xs.xmin) / xs.res;
  y = (xs.ymax - y) / xs.res;
  if ((int)y >= xs.nrows - 1 || (int)x >= xs.ncols - 1) {
    if ((int)y >= xs.nrows - 1) y = xs.nrows - 1;
    if ((int)x >= xs.ncols - 1) x = xs.ncols - 1;
    *zx = xs.at((int)y - 1, (int)x - 1);
    *zy = ys.at((int)y - 1, (int)x - 1);
  } else {
    *zx = bilinear(x, y, xs);
    *zy = bilinear(x, y, ys);
  }
}

// Get_curvature, GisApi.C:764-833
void get_curvature(const Raster& xs, const Raster& ys, double x, double y, double* cx,
                   double* cy) {
  int row = (int)((xs.ymax - y) / xs.res);
  if (row == 0) row++;
  if (row == xs.nrows - 1) row--;
  int col = (int)((x - xs.xmin) / xs.res);
  if (col == 0) col++;
  if (col == xs.ncols - 1) col--;
  if (col >= (xs.ncols - 3) || row >= xs.nrows - 3) {
    *cx = ((xs.at(row - 1, col + 1) - xs.at(row - 1, col - 1)) +
           2 * (xs.at(row, col + 1) - xs.at(row, col - 1)) +
           (xs.at(row + 1, col + 1) - xs.at(row + 1, col - 1))) / (8 * xs.res);
    *cy = ((ys.at(row - 1, col - 1) - ys.at(row + 1, col - 1)) +
           2 * (ys.at(row - 1, col) - ys.at(row + 1, col)) +
           (ys.at(row - 1, col + 1) - ys.at(row + 1, col + 1))) / (8 * xs.res);
    return;
  }
  Raster xx{2, 2, 0, 0, 1, std::vector<float>(4)}, yy{2, 2, 0, 0, 1, std::vector<float>(4)};
  const float res8 = 8 * (float)xs.res;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      xx.at(i, j) = ((xs.at(row + i - 1, col + j + 1) - xs.at(row + i - 1, col + j - 1)) +
                     2 * (xs.at(row + i, col + j + 1) - xs.at(row + i, col + j - 1)) +
                     (xs.at(row + i + 1, col + j + 1) - xs.at(row + i + 1, col + j - 1))) / res8;
      yy.at(i, j) = ((ys.at(row + i - 1, col + j - 1) - ys.at(row + i + 1, col + j - 1)) +
                     2 * (ys.at(row + i - 1, col + j) - ys.at(row + i + 1, col + j)) +
                     (ys.at(row + i - 1, col + j + 1) - ys.at(row + i + 1, col + j + 1))) / res8;
    }
  double x1 = x - xs.xmin - xs.res * (double)col;
  double y1 = xs.ymax - xs.res * (double)row - y;
  x1 = x1 / xs.res;
  y1 = y1 / xs.res;
  *cx = bilinear(x1, y1, xx);
  *cy = bilinear(x1, y1, yy);
}

// calc_gravity_vector, element2.C:3545-3564: gravity of the surface with
// slope zeta = (dz/dx, dz/dy)
void gravity_of(const double zeta[2], double g[3]) {
  const double max_slope = std::sqrt(zeta[0] * zeta[0] + zeta[1] * zeta[1]);
  const double max_angle = std::atan(max_slope);
  const double dsg = 9.8 * std::sin(max_angle);
  if (std::fabs(dsg) > kTiny) {
    g[0] = -dsg * zeta[0] / max_slope;
    g[1] = -dsg * zeta[1] / max_slope;
    g[2] = 9.8 * std::cos(max_angle);
  } else {
    g[0] = 0; g[1] = 0; g[2] = 9.8;
  }
}

// splitmix64: deterministic, seedable, platform independent
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double sym() { return 2.0 * uniform() - 1.0; }
};

}  // namespace

extern "C" int tg_host_dem_plane_runout(int nrows, int ncols, double xmin, double ymax,
                                        double res, double slope_deg, double x_toe,
                                        double blend, float* elev) {
  if (nrows < 3 || ncols < 3 || res <= 0 || !elev || blend < 0) return -1;
  const double t = std::tan(slope_deg * M_PI / 180.0);
  const double a = x_toe - 0.5 * blend, b = x_toe + 0.5 * blend;
  // z'(x) = -t * s(x), s = 1 (x<a), 0 (x>b), (1+cos(pi (x-a)/blend))/2 between;
  // z(x) = integral_x^inf t*s: C1 and exactly planar away from the blend.
  auto z = [&](double x) {
    double tail = 0.5 * blend;  // integral of s over [a,b]
    if (x >= b) return 0.0;
    if (x >= a) {
      double u = (x - a) / blend;
      // integral_x^b s = blend*( (1-u)/2 - sin(pi u)/(2 pi) )
      return t * blend * ((1.0 - u) * 0.5 - std::sin(M_PI * u) / (2.0 * M_PI));
    }
    return t * ((a - x) + tail);
  };
  for (int r = 0; r < nrows; r++) {
    double y = ymax - r * res;
    (void)y;
    for (int c = 0; c < ncols; c++) elev[(size_t)r * ncols + c] = (float)z(xmin + c * res);
  }
  return 0;
}

extern "C" int tg_host_dem_diamond_square(int n, double relief, double H, uint64_t seed,
                                          float* elev) {
  if (n < 3 || ((n - 1) & (n - 2)) != 0 || !elev || relief <= 0) return -1;
  std::vector<double> z((size_t)n * n, 0.0);
  auto Z = [&](int r, int c) -> double& { return z[(size_t)r * n + c]; };
  Rng rng(seed);
  Z(0, 0) = rng.sym(); Z(0, n - 1) = rng.sym(); Z(n - 1, 0) = rng.sym(); Z(n - 1, n - 1) = rng.sym();
  double amp = 1.0;
  const double fall = std::pow(2.0, -H);
  for (int step = n - 1; step > 1; step /= 2) {
    int half = step / 2;
    for (int r = half; r < n; r += step)          // diamond
      for (int c = half; c < n; c += step)
        Z(r, c) = 0.25 * (Z(r - half, c - half) + Z(r - half, c + half) + Z(r + half, c - half) +
                          Z(r + half, c + half)) + amp * rng.sym();
    for (int r = 0; r < n; r += half)              // square
      for (int c = (r / half % 2 == 0) ? half : 0; c < n; c += step) {
        double s = 0; int k = 0;
        if (r - half >= 0) { s += Z(r - half, c); k++; }
        if (r + half < n) { s += Z(r + half, c); k++; }
        if (c - half >= 0) { s += Z(r, c - half); k++; }
        if (c + half < n) { s += Z(r, c + half); k++; }
        Z(r, c) = s / k + amp * rng.sym();
      }
    amp *= fall;
  }
  double lo = z[0], hi = z[0];
  for (double v : z) { lo = std::fmin(lo, v); hi = std::fmax(hi, v); }
  const double sc = relief / (hi - lo);
  for (size_t i = 0; i < z.size(); i++) elev[i] = (float)((z[i] - lo) * sc);
  return 0;
}

extern "C" int tg_host_static_coefs(const tg_dem* dem, int nx, int ny, double dx, double dy,
                                    double x0, double y0, double length_scale,
                                    double gravity_scale, double* grav, double* dgrav,
                                    double* curv) {
  if (!dem || !dem->elev || nx < 1 || ny < 1 || !grav || !dgrav || !curv) return -1;
  Raster e{dem->nrows, dem->ncols, dem->xmin, dem->ymax, dem->res,
           std::vector<float>(dem->elev, dem->elev + (size_t)dem->nrows * dem->ncols)};
  // every cell centre must fall inside the raster (Get_slope's window test)
  double xmax = dem->xmin + (dem->ncols - 1) * dem->res, ymin = dem->ymax - (dem->nrows - 1) * dem->res;
  if (x0 * length_scale < dem->xmin || (x0 + (nx - 1) * dx) * length_scale > xmax ||
      y0 * length_scale < ymin || (y0 + (ny - 1) * dy) * length_scale > dem->ymax)
    return -2;
  Raster xs, ys;
  horn_slopes(e, xs, ys);
  const size_t n = (size_t)nx * ny;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      size_t c = (size_t)j * nx + i;
      double X = (x0 + i * dx) * length_scale, Y = (y0 + j * dy) * length_scale;
      double zeta[2], cv[2];
      get_slope(xs, ys, X, Y, &zeta[0], &zeta[1]);
      get_curvature(xs, ys, X, Y, &cv[0], &cv[1]);
      curv[c] = cv[0] * length_scale;
      curv[n + c] = cv[1] * length_scale;
      double g[3];
      gravity_of(zeta, g);
      for (int k = 0; k < 3; k++) grav[k * n + c] = g[k] / gravity_scale;
    }
  // calc_d_gravity, element2.C:3574-3653 (one-sided where a neighbour is missing)
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      size_t c = (size_t)j * nx + i;
      const double* gz = grav + 2 * n;
      for (int dir = 0; dir < 2; dir++) {
        bool hp = dir == 0 ? i + 1 < nx : j + 1 < ny;
        bool hm = dir == 0 ? i > 0 : j > 0;
        size_t ep = dir == 0 ? c + 1 : c + nx, em = dir == 0 ? c - 1 : c - nx;
        double crd = dir == 0 ? x0 + i * dx : y0 + j * dy;
        double step = dir == 0 ? dx : dy;
        double out;
        if (hp && hm) {
          double dxp = (crd + step) - crd, dxm = crd - (crd - step);
          double dp = (gz[ep] - gz[c]) / dxp, dm = (gz[c] - gz[em]) / dxm;
          out = (dp * dxm + dm * dxp) / (dxm + dxp);
        } else if (hm) {
          double dxm = crd - (crd - step);
          out = (gz[c] - gz[em]) / dxm;
        } else if (hp) {
          double dxp = (crd + step) - crd;
          out = (gz[ep] - gz[c]) / dxp;
        } else {
          out = 0;
        }
        dgrav[dir * n + c] = out;
      }
    }
  return 0;
}

extern "C" int tg_host_leaf_static_coefs(const tg_dem* dem, int n, const double* xc,
                                         double length_scale, double gravity_scale,
                                         double* grav, double* curv) {
  if (!dem || !dem->elev || n < 0 || (n > 0 && (!xc || !grav || !curv))) return -1;
  Raster e{dem->nrows, dem->ncols, dem->xmin, dem->ymax, dem->res,
           std::vector<float>(dem->elev, dem->elev + (size_t)dem->nrows * dem->ncols)};
  const double xmax = dem->xmin + (dem->ncols - 1) * dem->res;
  const double ymin = dem->ymax - (dem->nrows - 1) * dem->res;
  for (int c = 0; c < n; c++) {
    const double X = xc[2 * c] * length_scale, Y = xc[2 * c + 1] * length_scale;
    if (X < dem->xmin || X > xmax || Y < ymin || Y > dem->ymax) return -2;
  }
  Raster xs, ys;
  horn_slopes(e, xs, ys);
  for (int c = 0; c < n; c++) {
    const double X = xc[2 * c] * length_scale, Y = xc[2 * c + 1] * length_scale;
    double zeta[2], cv[2], g[3];
    get_slope(xs, ys, X, Y, &zeta[0], &zeta[1]);
    get_curvature(xs, ys, X, Y, &cv[0], &cv[1]);
    curv[c] = cv[0] * length_scale;
    curv[(size_t)n + c] = cv[1] * length_scale;
    gravity_of(zeta, g);
    for (int k = 0; k < 3; k++) grav[(size_t)k * n + c] = g[k] / gravity_scale;
  }
  return 0;
}

extern "C" int tg_host_piles(int nx, int ny, double dx, double dy, double x0, double y0,
                             int npiles, const double* xcen, const double* ycen,
                             const double* major, const double* minor, const double* rot_deg,
                             const double* height, double* h, double* hu, double* hv) {
  if (nx < 1 || ny < 1 || npiles < 0 || !h || !hu || !hv) return -1;
  std::vector<double> cr(npiles), sr(npiles);
  for (int p = 0; p < npiles; p++) {
    cr[p] = std::cos(rot_deg[p] * M_PI / 180.0);
    sr[p] = std::sin(rot_deg[p] * M_PI / 180.0);
  }
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      double xc = x0 + i * dx, yc = y0 + j * dy;
      // node order: corners 0-3, edges 4-7, centre 8 (weights 1,2,4)
      const double nxs[9] = {xc - 0.5 * dx, xc + 0.5 * dx, xc + 0.5 * dx, xc - 0.5 * dx,
                             xc, xc + 0.5 * dx, xc, xc - 0.5 * dx, xc};
      const double nys[9] = {yc - 0.5 * dy, yc - 0.5 * dy, yc + 0.5 * dy, yc + 0.5 * dy,
                             yc - 0.5 * dy, yc, yc + 0.5 * dy, yc, yc};
      double nh[9];
      for (int q = 0; q < 9; q++) {
        nh[q] = 0.0;
        for (int p = 0; p < npiles; p++) {
          double ma = nxs[q] - xcen[p], mi = nys[q] - ycen[p];
          double sw = (ma * cr[p] + mi * sr[p]) / major[p];
          mi = (-ma * sr[p] + mi * cr[p]) / minor[p];
          ma = sw;
          double ht = height[p] * (1. - ma * ma - mi * mi);
          ht = (ht >= 0.0) ? ht : 0.0;
          if (nh[q] < ht) nh[q] = ht;
        }
      }
      size_t c = (size_t)j * nx + i;
      h[c] = (nh[0] + nh[1] + nh[2] + nh[3] + 2.0 * (nh[4] + nh[5] + nh[6] + nh[7]) +
              4.0 * nh[8]) / 16.0;
      hu[c] = 0.0;
      hv[c] = 0.0;
    }
  return 0;
}

extern "C" int tg_host_box_weights(int nx, int ny, double dx, double dy, double x0, double y0,
                                   double xmin, double xmax, double ymin, double ymax,
                                   double* w) {
  if (nx < 1 || ny < 1 || !w) return -1;
  const size_t n = (size_t)nx * ny;
  std::memset(w, 0, sizeof(double) * 3 * n);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      double x = x0 + i * dx, y = y0 + j * dy;
      if (x > xmin && x < xmax && y > ymin && y < ymax) w[(size_t)j * nx + i] = dx * dy;
    }
  return 0;
}

// calc_d_gravity (element2.C:3574-3653) on a flattened leaf table: the
// neighbour of side xp / xm / yp / ym is the leaf in slot 1 / 3 / 2 / 0
// (positive_x_side = 1) -- on a side with two finer leaves, the one of
// slot s (neighbor[s]), never slot s + 4 -- and the spacing is the
// difference of the centres along the direction; a side without neighbour
// (-1: domain boundary) makes the difference one-sided, none on either side
// gives 0.
extern "C" int tg_host_leaf_d_gravity(int n, const int32_t* nbr, const double* xc,
                                      const double* gz, double* dgrav) {
  if (n < 0 || (n > 0 && (!nbr || !xc || !gz || !dgrav))) return -1;
  for (int e = 0; e < n; e++)
    for (int s = 0; s < 8; s++)
      if (nbr[8 * e + s] < -1 || nbr[8 * e + s] >= n) return -1;
  static const int kP[2] = {1, 2}, kM[2] = {3, 0};  // xp, yp; xm, ym
  for (int e = 0; e < n; e++)
    for (int d = 0; d < 2; d++) {
      const int ep = nbr[8 * e + kP[d]], em = nbr[8 * e + kM[d]];
      const double c = xc[2 * e + d];
      double out;
      if (ep >= 0 && em >= 0) {
        const double dxp = xc[2 * ep + d] - c, dxm = c - xc[2 * em + d];
        const double dp = (gz[ep] - gz[e]) / dxp, dm = (gz[e] - gz[em]) / dxm;
        out = (dp * dxm + dm * dxp) / (dxm + dxp);  // weighted average
      } else if (em >= 0) {
        const double dxm = c - xc[2 * em + d];
        out = (gz[e] - gz[em]) / dxm;
      } else if (ep >= 0) {
        const double dxp = xc[2 * ep + d] - c;
        out = (gz[ep] - gz[e]) / dxp;
      } else {
        out = 0;  // no neighbour on either side: flat ground
      }
      dgrav[(size_t)d * n + e] = out;
    }
  return 0;
}
