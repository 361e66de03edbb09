// Host-side flattening of an adaptive quadtree into the leaf tables the
// reference's traversal produces: Hilbert keys of the leaf centres, the
// hash-bucket traversal order of the element table, and the 8-slot
// neighbour / generation / neighbour-kind tables (two neighbours on a side
// that borders finer leaves, one otherwise).
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "../../include/tg_host.h"

namespace {

// Inverse 2D Hilbert map over [0, 2^32-1]^2 to a 64-bit key, two bits per
// level, most significant level first; the orientation state is a pair of
// axis positions (order) and a reflection mask, as in the published Zoltan
// BSFC formulation (hsfc2d, hilbert.C:76-145).  The 2-bit Gray code inverse
// maps quadrant bits g -> curve position: 0->0, 1->1, 3->2, 2->3.
void hilbert2(uint32_t cx, uint32_t cy, uint32_t key[2]) {
  static const unsigned char gray_inv[4] = {0, 1, 3, 2};
  unsigned ox = 0, oy = 1;  // bit position of the x and y bits in the quadrant code
  unsigned refl = 0;
  uint64_t k = 0;
  for (int lev = 31; lev >= 0; lev--) {
    const unsigned q = (((cx >> lev) & 1u) << ox) | (((cy >> lev) & 1u) << oy);
    const unsigned c = gray_inv[refl ^ q];
    k = (k << 2) | c;
    if (c == 0 || c == 3) {  // first and last sub-square: transpose the axes
      if (c == 3) refl ^= 3u;
      const unsigned t = ox;
      ox = oy;
      oy = t;
    }
  }
  key[0] = (uint32_t)(k >> 32);
  key[1] = (uint32_t)k;
}

// Normalised coordinate -> integer coordinate (fhsfc2d_, hilbert.C:267-278):
// scaled by 2^32-1 in FP64, clamped, truncated.
uint32_t scale_coord(double u) {
  const double imax = 4294967295.0;
  const double c = (u <= 0.) ? 0. : imax * u;
  return (uint32_t)((imax < c) ? imax : c);
}

uint64_t cell_code(int lev, int i, int j) {
  return ((uint64_t)(uint32_t)lev << 58) ^ ((uint64_t)(uint32_t)i << 29) ^ (uint64_t)(uint32_t)j;
}

// Element-table insertion into a bucket's list (HashTable::addElement,
// hashtab2.C:185-217): walk past entries with a smaller first key word; if
// the entry reached has an equal first word, keep walking while the second
// word is larger (without re-checking the first word).
void bucket_insert(std::vector<int>& list, const uint32_t* keys, int e) {
  const uint32_t k0 = keys[2 * e], k1 = keys[2 * e + 1];
  size_t p = 0;
  while (p < list.size() && k0 > keys[2 * list[p]]) p++;
  if (p < list.size() && k0 == keys[2 * list[p]])
    while (p < list.size() && k1 > keys[2 * list[p] + 1]) p++;
  list.insert(list.begin() + (long)p, e);
}

}  // namespace

extern "C" {

int tg_host_hsfc2d(int n, const double* norm_xy, uint32_t* keys) {
  if (n < 0 || (n > 0 && (!norm_xy || !keys))) return -1;
  for (int c = 0; c < n; c++)
    hilbert2(scale_coord(norm_xy[2 * c]), scale_coord(norm_xy[2 * c + 1]), keys + 2 * c);
  return 0;
}

int tg_host_hash_bucket(int n, const uint32_t* keys, double keyrange0, double keyrange1,
                        int nbuckets, int32_t* bucket) {
  if (n < 0 || nbuckets < 1 || (n > 0 && (!keys || !bucket))) return -1;
  const double hc = 8.0 * nbuckets / (keyrange0 * keyrange1 + keyrange1);
  for (int c = 0; c < n; c++) {
    const double v = (keys[2 * c] * keyrange1 + keys[2 * c + 1]) * hc + 0.5;
    bucket[c] = ((int)v) % nbuckets;
  }
  return 0;
}

int tg_host_traversal(int n, const uint32_t* keys, double keyrange0, double keyrange1,
                      int nbuckets, int32_t* order) {
  if (n < 0 || nbuckets < 1 || (n > 0 && (!keys || !order))) return -1;
  std::vector<int32_t> bucket(n);
  tg_host_hash_bucket(n, keys, keyrange0, keyrange1, nbuckets, bucket.data());
  std::vector<std::vector<int>> lists(nbuckets);
  for (int e = 0; e < n; e++) bucket_insert(lists[bucket[e]], keys, e);
  int p = 0;
  for (int b = 0; b < nbuckets; b++)
    for (int e : lists[b]) order[p++] = e;
  return 0;
}

int tg_host_flatten(int nx0, int ny0, double xmin, double ymin, double dx0, double dy0, int n,
                    const int32_t* lev, const int32_t* ix, const int32_t* iy, int nbuckets,
                    tg_leaf_table* out) {
  if (nx0 < 1 || ny0 < 1 || n < 1 || nbuckets < 1 || !lev || !ix || !iy || !out ||
      !(dx0 > 0.) || !(dy0 > 0.))
    return -1;
  if (!out->order || !out->key || !out->gen || !out->nbr || !out->nkind || !out->ngen ||
      !out->dx || !out->xc)
    return -1;
  // leaf lookup by (level, i, j); reject overlaps, levels out of range and
  // cells outside the base grid
  std::unordered_map<uint64_t, int> leaf;
  leaf.reserve((size_t)n * 2);
  double area = 0.;
  for (int e = 0; e < n; e++) {
    const int L = lev[e];
    if (L < 0 || L > 24) return -2;
    const long long W = (long long)nx0 << L, H = (long long)ny0 << L;
    if (ix[e] < 0 || iy[e] < 0 || ix[e] >= W || iy[e] >= H) return -2;
    if (!leaf.emplace(cell_code(L, ix[e], iy[e]), e).second) return -2;
    area += std::ldexp(1.0, -2 * L);
  }
  auto find = [&](int L, long long i, long long j) -> int {
    if (L < 0 || i < 0 || j < 0 || i >= ((long long)nx0 << L) || j >= ((long long)ny0 << L))
      return -1;
    auto it = leaf.find(cell_code(L, (int)i, (int)j));
    return it == leaf.end() ? -1 : it->second;
  };
  // ancestors must not be leaves (overlap of a leaf and its descendant)
  for (int e = 0; e < n; e++)
    for (int L = lev[e] - 1, i = ix[e] >> 1, j = iy[e] >> 1; L >= 0; L--, i >>= 1, j >>= 1)
      if (find(L, i, j) >= 0) return -2;
  if (area != (double)nx0 * ny0) return -3;  // leaves do not tile the base grid

  const double xmax = xmin + nx0 * dx0, ymax = ymin + ny0 * dy0;
  std::vector<double> xc(2 * (size_t)n), dxl(2 * (size_t)n);
  std::vector<uint32_t> keys(2 * (size_t)n);
  for (int e = 0; e < n; e++) {
    const double hx = std::ldexp(dx0, -lev[e]), hy = std::ldexp(dy0, -lev[e]);
    dxl[2 * e] = hx;
    dxl[2 * e + 1] = hy;
    xc[2 * e] = xmin + (ix[e] + 0.5) * hx;
    xc[2 * e + 1] = ymin + (iy[e] + 0.5) * hy;
    const double nxy[2] = {(xc[2 * e] - xmin) / (xmax - xmin), (xc[2 * e + 1] - ymin) / (ymax - ymin)};
    tg_host_hsfc2d(1, nxy, keys.data() + 2 * e);
  }
  // element table of one process: key ranges (2^32 + 1) for both words
  // (preprocess.C:561-562 with np = 1)
  const double kr = 4294967296.0 + 1.0;
  tg_host_traversal(n, keys.data(), kr, kr, nbuckets, out->order);
  std::vector<int32_t> pos(n);
  for (int p = 0; p < n; p++) pos[out->order[p]] = p;

  // Neighbour slots: side s in {0 south, 1 east, 2 north, 3 west} (corners
  // 0 SW, 1 SE, 2 NE, 3 NW; side s runs from corner s to corner s+1).  With
  // two finer neighbours, slot s holds the one next to corner s and slot
  // s+4 the one next to corner s+1 (son construction, refine2.C:589-700).
  static const int dI[4] = {0, 1, 0, -1}, dJ[4] = {-1, 0, 1, 0};
  for (int e = 0; e < n; e++) {
    const int q = pos[e], L = lev[e];
    out->key[2 * q] = keys[2 * e];
    out->key[2 * q + 1] = keys[2 * e + 1];
    out->gen[q] = L;
    out->dx[2 * q] = dxl[2 * e];
    out->dx[2 * q + 1] = dxl[2 * e + 1];
    out->xc[2 * q] = xc[2 * e];
    out->xc[2 * q + 1] = xc[2 * e + 1];
    for (int s = 0; s < 4; s++) {
      const long long i = (long long)ix[e] + dI[s], j = (long long)iy[e] + dJ[s];
      int32_t* nb = out->nbr + 8 * (size_t)q;
      int32_t* nk = out->nkind + 8 * (size_t)q;
      int32_t* ng = out->ngen + 8 * (size_t)q;
      if (i < 0 || j < 0 || i >= ((long long)nx0 << L) || j >= ((long long)ny0 << L)) {
        nb[s] = nb[s + 4] = -1;  // domain boundary (neigh_proc == -1)
        nk[s] = -1;
        nk[s + 4] = -2;
        ng[s] = ng[s + 4] = L;
        continue;
      }
      int f = find(L, i, j);
      if (f < 0) f = find(L - 1, i >> 1, j >> 1);
      if (f >= 0) {  // one neighbour on this side, same or coarser level
        nb[s] = nb[s + 4] = pos[f];
        nk[s] = 0;
        nk[s + 4] = -2;
        ng[s] = ng[s + 4] = lev[f];
        continue;
      }
      // two finer neighbours: the level-(L+1) children of (i, j) touching side s
      const long long ci = 2 * i, cj = 2 * j;
      long long a[2][2];  // [near corner s, near corner s+1][i, j]
      switch (s) {
        case 0: a[0][0] = ci;     a[0][1] = cj + 1; a[1][0] = ci + 1; a[1][1] = cj + 1; break;
        case 1: a[0][0] = ci;     a[0][1] = cj;     a[1][0] = ci;     a[1][1] = cj + 1; break;
        case 2: a[0][0] = ci + 1; a[0][1] = cj;     a[1][0] = ci;     a[1][1] = cj;     break;
        default: a[0][0] = ci + 1; a[0][1] = cj + 1; a[1][0] = ci + 1; a[1][1] = cj;    break;
      }
      const int g0 = find(L + 1, a[0][0], a[0][1]), g1 = find(L + 1, a[1][0], a[1][1]);
      if (g0 < 0 || g1 < 0) return -4;  // more than one level of difference (2:1 rule)
      nb[s] = pos[g0];
      nb[s + 4] = pos[g1];
      nk[s] = nk[s + 4] = 0;
      ng[s] = ng[s + 4] = L + 1;
    }
  }
  out->n = n;
  return 0;
}

}  // extern "C"
