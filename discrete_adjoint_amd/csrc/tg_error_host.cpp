// Interpolation stencils of the error grid: the reference's once-refined
// copy of the mesh (uinform_refine, uniform_refine.C:24-130), on which
// bilinear_interp (bilinear_interp.C:126-289) reconstructs the state, the
// previous state and the adjoint of every son from the sons of neighbouring
// leaves.  Host code, once per mesh: for every error element it picks the
// four elements of its interpolation stencil exactly as bilinear_interp does
// (the twelve side neighbours of a family of four sons through
// Element::get_side_neighbor, element2.C:4224-4330, and the coincidence
// cases), turns them into the points bilinear_interp_elem interpolates
// (the centre of each element's father, from its own centre and which_son,
// with the values its father carries; points closer than 1e-12 merged,
// bilinear_interp.C:55-106, 427-459) and classifies the rule: one point
// (the element keeps its own values), two (linear along x or y, sorted as
// NeighLess), three (barycentric) or four (the 4x4 bilinear fit).
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/tg_host.h"

namespace {

// positive_x_side = 1 on generated grids: x+ is slot 1, y+ slot 2, x- slot 3,
// y- slot 0 (the second neighbour of side s in slot s + 4)
constexpr int XP = 1, YP = 2, XM = 3, YM = 0;
constexpr int INIT = -1;  // neigh_proc of a domain-boundary side

struct Grid {
  const int32_t* nbr;
  const int32_t* np;
};

// Element::get_side_neighbor<T> (element2.C:4224-4330): -1 stands for the
// NULL a failed lookup returns (a boundary side's neighbour key).
int side_nb(const Grid& g, int e, int side) {
  if (e < 0) return -1;
  const int32_t* nb = g.nbr + 8 * (size_t)e;
  const int32_t* np = g.np + 8 * (size_t)e;
  switch (side) {
    case 0: return np[XP] != INIT ? nb[XP] : -1;
    case 1: return np[YP] != INIT ? nb[YP] : -1;
    case 2: return np[XM] != INIT ? nb[XM] : -1;
    case 3: return np[YM] != INIT ? nb[YM] : -1;
    case 4: return np[XP + 4] < 0 ? nb[XP] : nb[XP + 4];
    case 5: return np[YP] < 0 ? nb[YP] : nb[YP + 4];
    case 6: return np[XM] < 0 ? nb[XM] : nb[XM + 4];
    case 7: return np[YM] < INIT ? nb[YM] : nb[YM + 4];
    case 8: return side_nb(g, side_nb(g, e, 1), 0);
    case 9: return side_nb(g, side_nb(g, e, 2), 1);
    case 10: return side_nb(g, side_nb(g, e, 3), 2);
    case 11: return side_nb(g, side_nb(g, e, 0), 3);
    default: return -1;
  }
}

struct Pt {
  double x, y;
  int leaf;
};

bool same(const Pt& a, const Pt& b) {
  return std::fabs(a.x - b.x) < 1e-12 && std::fabs(a.y - b.y) < 1e-12;
}
// NeighLess (bilinear_interp.C:114-121)
bool less(const Pt& a, const Pt& b) {
  return (a.x < b.x && std::fabs(a.x - b.x) > 1e-12) ||
         (std::fabs(a.x - b.x) < 1e-12 && a.y < b.y);
}

}  // namespace

extern "C" int tg_host_error_stencil(int n, const int32_t* nbr, const int32_t* nproc,
                                     const double* cdx, const double* xc,
                                     const int32_t* which_son, const int32_t* father,
                                     int nfather, int32_t* kind, int32_t* pts, double* pxy) {
  if (n < 0 || nfather < 0 || (n > 0 && (!nbr || !nproc || !cdx || !xc || !which_son ||
                                         !father || !kind || !pts || !pxy)))
    return -1;
  // the four sons of every father, by which_son
  std::vector<int32_t> sons(4 * (size_t)nfather, -1);
  for (int e = 0; e < n; e++) {
    const int f = father[e], w = which_son[e];
    if (f < 0 || f >= nfather || w < 0 || w > 3 || sons[4 * (size_t)f + w] >= 0) return -2;
    sons[4 * (size_t)f + w] = e;
  }
  for (int f = 0; f < nfather; f++)
    for (int w = 0; w < 4; w++)
      if (sons[4 * (size_t)f + w] < 0) return -2;
  const Grid g{nbr, nproc};
  // the centre of element e's father (Neighbor, bilinear_interp.C:55-90)
  auto point = [&](int e) {
    const double hx = cdx[2 * (size_t)e], hy = cdx[2 * (size_t)e + 1];
    const double x = xc[2 * (size_t)e], y = xc[2 * (size_t)e + 1];
    Pt p;
    switch (which_son[e]) {
      case 0: p.x = x + .5 * hx; p.y = y + .5 * hy; break;
      case 1: p.x = x - .5 * hx; p.y = y + .5 * hy; break;
      case 2: p.x = x - .5 * hx; p.y = y - .5 * hy; break;
      default: p.x = x + .5 * hx; p.y = y - .5 * hy; break;
    }
    p.leaf = father[e];
    return p;
  };
  for (int f = 0; f < nfather; f++) {
    const int32_t* s = &sons[4 * (size_t)f];
    int nb[12];
    nb[3] = side_nb(g, s[0], 3);
    nb[6] = side_nb(g, s[0], 6);
    nb[10] = side_nb(g, s[0], 10);
    nb[0] = side_nb(g, s[1], 0);
    nb[7] = side_nb(g, s[1], 7);
    nb[11] = side_nb(g, s[1], 11);
    nb[8] = side_nb(g, s[2], 8);
    nb[4] = side_nb(g, s[2], 4);
    nb[1] = side_nb(g, s[2], 1);
    nb[5] = side_nb(g, s[3], 5);
    nb[9] = side_nb(g, s[3], 9);
    nb[2] = side_nb(g, s[3], 2);
    auto eq = [&](int a, int b) { return nb[a] >= 0 && nb[b] >= 0 && nb[a] == nb[b]; };
    for (int j = 0; j < 4; j++) {
      const int e = s[j];
      const int32_t* np = nproc + 8 * (size_t)e;
      int st[4];  // elem11, elem21, elem12, elem22 (bilinear_interp.C:178-279)
      auto set = [&](int a, int b, int c, int d) { st[0] = a; st[1] = b; st[2] = c; st[3] = d; };
      switch (j) {
        case 0:
          if (eq(6, 10)) {
            if (np[XM + 4] < 0) set(nb[10], nb[3], nb[9], e);
            else set(nb[10], nb[3], nb[2], e);
          } else if (eq(3, 10)) {
            if (np[YM + 4] < 0) set(nb[10], nb[11], nb[6], e);
            else set(nb[10], nb[7], nb[6], e);
          } else {
            set(nb[10], nb[3], nb[6], e);
          }
          break;
        case 1:
          if (eq(11, 0)) {
            if (np[XP + 4] < 0) set(nb[7], nb[11], e, nb[8]);
            else set(nb[7], nb[11], e, nb[4]);
          } else if (eq(7, 11)) {
            if (np[YM + 4] < 0) set(nb[10], nb[11], e, nb[0]);
            else set(nb[3], nb[11], e, nb[0]);
          } else {
            set(nb[7], nb[11], e, nb[0]);
          }
          break;
        case 2:
          if (eq(4, 8)) {
            if (np[XP + 4] < 0) set(e, nb[11], nb[1], nb[8]);
            else set(e, nb[0], nb[1], nb[8]);
          } else if (eq(8, 1)) {
            if (np[YP + 4] < 0) set(e, nb[4], nb[9], nb[8]);
            else set(e, nb[4], nb[5], nb[8]);
          } else {
            set(e, nb[4], nb[1], nb[8]);
          }
          break;
        default:
          if (eq(9, 2)) {
            if (np[XM + 4] < 0) set(nb[10], e, nb[9], nb[5]);
            else set(nb[6], e, nb[9], nb[5]);
          } else if (eq(9, 5)) {
            if (np[YP + 4] < 0) set(nb[2], e, nb[9], nb[8]);
            else set(nb[2], e, nb[9], nb[1]);
          } else {
            set(nb[2], e, nb[9], nb[5]);
          }
          break;
      }
      // distinct points in stencil order (bilinear_interp_elem, :439-459)
      Pt u[4];
      int m = 0;
      for (int k = 0; k < 4; k++) {
        if (st[k] < 0) continue;
        const Pt p = point(st[k]);
        bool dup = false;
        for (int q = 0; q < m && !dup; q++) dup = same(p, u[q]);
        if (!dup) u[m++] = p;
      }
      int kd;
      switch (m) {
        case 1: kd = TG_ERR_INTERP_COPY; break;
        case 2:
          if (less(u[1], u[0])) {
            const Pt t = u[0];
            u[0] = u[1];
            u[1] = t;
          }
          kd = (u[1].x - u[0].x < 1e-12) ? TG_ERR_INTERP_LINEAR_Y : TG_ERR_INTERP_LINEAR_X;
          break;
        case 3: kd = TG_ERR_INTERP_BARYCENTRIC; break;
        case 4: kd = TG_ERR_INTERP_BILINEAR; break;
        default: return -3;  // no point: cannot happen, the element is its own
      }
      kind[e] = kd;
      for (int k = 0; k < 4; k++) {
        pts[4 * (size_t)e + k] = k < m ? u[k].leaf : -1;
        pxy[8 * (size_t)e + 2 * k] = k < m ? u[k].x : 0.;
        pxy[8 * (size_t)e + 2 * k + 1] = k < m ? u[k].y : 0.;
      }
    }
  }
  return 0;
}
