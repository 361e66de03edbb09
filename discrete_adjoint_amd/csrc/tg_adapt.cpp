// Mesh adaptation between steps: one pass of H_adapt (refinement along the
// pile contours, 2:1 cascade, buffer layers; hadpt.C:50-323) followed by
// unrefine (unrefine.C:54-352), on a quadtree given by its leaves, with the
// state transfer of the reference (a son copies its father,
// element2.C:246-250; a new father is the 0.25-weighted sum of its sons,
// element2.C:527-532) and its transpose for the adjoint (dualmesh.C:708,
// 986-993: a son takes 0.25 of its father's adjoint, a father the sum of its
// sons').  Host code: the tree changes every few steps and feeds
// tg_host_flatten / tg_set_mesh.
#include "../../include/tg_host.h"

#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace {

// adapted flags (constant.h:77-84); only the ones a pass produces
constexpr int NOTRECADAPTED = 1, NEWFATHER = 2, NEWSON = 3, BUFFER = 4;

struct Node {
  double U[3];
  double err;
  int flag;
};

inline uint64_t ckey(int lev, int64_t ix, int64_t iy) {
  return ((uint64_t)lev << 58) | ((uint64_t)ix << 29) | (uint64_t)iy;
}

struct Tree {
  int nx0, ny0, max_level;
  std::unordered_map<uint64_t, Node> leaf;
  std::vector<tg_adapt_op> ops;

  bool inside(int lev, int64_t i, int64_t j) const {
    return i >= 0 && j >= 0 && i < ((int64_t)nx0 << lev) && j < ((int64_t)ny0 << lev);
  }
  const Node* at(int lev, int64_t i, int64_t j) const {
    auto it = leaf.find(ckey(lev, i, j));
    return it == leaf.end() ? nullptr : &it->second;
  }

  // Leaves across side s (0 S, 1 E, 2 N, 3 W) of leaf (lev, i, j): none on
  // the domain boundary, one of the same or the next coarser level, or the
  // two finer leaves along the side.  Returns -1 if the tree is not 2:1
  // balanced there.
  int side_nbrs(int lev, int64_t i, int64_t j, int s, int* nl, int64_t* ni, int64_t* nj) const {
    static const int DI[4] = {0, 1, 0, -1}, DJ[4] = {-1, 0, 1, 0};
    const int64_t a = i + DI[s], b = j + DJ[s];
    if (!inside(lev, a, b)) return 0;
    if (at(lev, a, b)) {
      nl[0] = lev; ni[0] = a; nj[0] = b;
      return 1;
    }
    if (lev > 0 && at(lev - 1, a >> 1, b >> 1)) {
      nl[0] = lev - 1; ni[0] = a >> 1; nj[0] = b >> 1;
      return 1;
    }
    // the two sons of (lev, a, b) that touch side s of the leaf
    int64_t si[2], sj[2];
    switch (s) {
      case 0: si[0] = 2 * a; si[1] = 2 * a + 1; sj[0] = sj[1] = 2 * b + 1; break;
      case 1: si[0] = si[1] = 2 * a; sj[0] = 2 * b; sj[1] = 2 * b + 1; break;
      case 2: si[0] = 2 * a + 1; si[1] = 2 * a; sj[0] = sj[1] = 2 * b; break;
      default: si[0] = si[1] = 2 * a + 1; sj[0] = 2 * b + 1; sj[1] = 2 * b; break;
    }
    for (int k = 0; k < 2; k++) {
      if (!at(lev + 1, si[k], sj[k])) return -1;
      nl[k] = lev + 1; ni[k] = si[k]; nj[k] = sj[k];
    }
    return 2;
  }

  // every leaf next to (lev, i, j), as the reference's neighbour slots with
  // neigh_proc >= 0
  int nbrs(int lev, int64_t i, int64_t j, const Node** out, int* out_lev) const {
    int cnt = 0;
    for (int s = 0; s < 4; s++) {
      int nl[2];
      int64_t ni[2], nj[2];
      const int m = side_nbrs(lev, i, j, s, nl, ni, nj);
      for (int k = 0; k < m; k++) {
        out[cnt] = at(nl[k], ni[k], nj[k]);
        out_lev[cnt++] = nl[k];
      }
    }
    return cnt;
  }

  // Splits leaf (lev, i, j) into its sons, first splitting every coarser
  // leaf across its sides so the mesh stays 2:1 balanced (refinewrapper's
  // triggered refinement, hadpt.C:325-372).  Sons copy the state and are
  // NEWSON.
  void refine(int lev, int64_t i, int64_t j) {
    if (!at(lev, i, j) || lev >= 62) return;
    for (int s = 0; s < 4; s++) {
      int nl[2];
      int64_t ni[2], nj[2];
      const int m = side_nbrs(lev, i, j, s, nl, ni, nj);
      if (m == 1 && nl[0] < lev) refine(nl[0], ni[0], nj[0]);
    }
    const Node f = *at(lev, i, j);
    leaf.erase(ckey(lev, i, j));
    static const int SI[4] = {0, 1, 1, 0}, SJ[4] = {0, 0, 1, 1};  // SW SE NE NW
    for (int k = 0; k < 4; k++) {
      Node s = f;
      for (int v = 0; v < 3; v++) s.U[v] = f.U[v] * 1.0;  // myfractionoffather = 1
      s.err = f.err * .5;
      s.flag = NEWSON;
      leaf[ckey(lev + 1, 2 * i + SI[k], 2 * j + SJ[k])] = s;
    }
    ops.push_back(tg_adapt_op{0, lev, (int32_t)i, (int32_t)j});
  }
};

// if_pile_boundary (element2.C:4020-4065): 2 inside the contour next to a
// leaf outside it, 1 outside next to a leaf inside, else 0
int pile_boundary(const Tree& T, int lev, int64_t i, int64_t j, double contour) {
  const Node* nb[8];
  int nlev[8];
  const int m = T.nbrs(lev, i, j, nb, nlev);
  const double h = T.at(lev, i, j)->U[0];
  if (h >= contour) {
    for (int k = 0; k < m; k++)
      if (nb[k]->U[0] < contour) return 2;
  } else {
    for (int k = 0; k < m; k++)
      if (nb[k]->U[0] >= contour) return 1;
  }
  return 0;
}

// if_first_buffer_boundary (element2.C:4136-4180) without mass sources
int first_buffer(const Tree& T, int lev, int64_t i, int64_t j, double contour) {
  const Node* self = T.at(lev, i, j);
  if (self->flag <= 0) return self->flag - 1;
  const Node* nb[8];
  int nlev[8];
  const int m = T.nbrs(lev, i, j, nb, nlev);
  bool buf = false;
  if (self->U[0] < contour) {
    for (int k = 0; k < m && !buf; k++) buf = nb[k]->U[0] >= contour;
  } else {
    for (int k = 0; k < m && !buf; k++) buf = nb[k]->U[0] < contour;
  }
  if (!buf) return 0;
  return (self->flag >= NEWSON || lev == T.max_level) ? 2 : 1;
}

// if_next_buffer_boundary (element2.C:4182-4220) without mass sources
int next_buffer(const Tree& T, int lev, int64_t i, int64_t j) {
  const Node* self = T.at(lev, i, j);
  if (self->flag <= 0) return self->flag - 1;
  bool nxt = false;
  if (self->flag != BUFFER) {
    const Node* nb[8];
    int nlev[8];
    const int m = T.nbrs(lev, i, j, nb, nlev);
    for (int k = 0; k < m && !nxt; k++)
      nxt = (nb[k]->flag == BUFFER) && (self->U[0] <= nb[k]->U[0]);
  }
  if (!nxt) return 0;
  return (self->flag >= NEWSON || lev == T.max_level) ? 2 : 1;
}

struct Cell {
  int lev;
  int64_t i, j;
};

// the current leaves in a fixed order: the input order, a refined leaf
// replaced by its sons (SW, SE, NE, NW, recursively); only refinements may
// have happened since the input
void ordered_leaves(const Tree& T, const std::vector<Cell>& base, std::vector<Cell>& out) {
  out.clear();
  std::vector<Cell> stack;
  static const int SI[4] = {0, 1, 1, 0}, SJ[4] = {0, 0, 1, 1};
  for (const Cell& c0 : base) {
    stack.push_back(c0);
    while (!stack.empty()) {
      const Cell c = stack.back();
      stack.pop_back();
      if (T.at(c.lev, c.i, c.j)) {
        out.push_back(c);
      } else if (c.lev < T.max_level) {
        // pushed in reverse so the sons come out SW, SE, NE, NW
        for (int k = 3; k >= 0; k--) stack.push_back(Cell{c.lev + 1, 2 * c.i + SI[k], 2 * c.j + SJ[k]});
      }
    }
  }
}

}  // namespace

extern "C" int tg_host_adapt(int nx0, int ny0, const tg_adapt_params* prm, int n,
                             const int32_t* lev, const int32_t* ix, const int32_t* iy,
                             const double* h, const double* hu, const double* hv,
                             const double* el_error, int cap, int* n_out, int32_t* lev_out,
                             int32_t* ix_out, int32_t* iy_out, double* h_out, double* hu_out,
                             double* hv_out, int op_cap, int* nops, tg_adapt_op* ops) {
  if (nx0 < 1 || ny0 < 1 || !prm || n < 1 || !lev || !ix || !iy || !h || !hu || !hv ||
      !n_out || !nops || prm->max_level < 0 || prm->max_level > 28 || prm->num_buffer_layer < 0)
    return -1;
  Tree T;
  T.nx0 = nx0; T.ny0 = ny0; T.max_level = prm->max_level;
  std::vector<Cell> base(n);
  for (int e = 0; e < n; e++) {
    if (lev[e] < 0 || lev[e] > prm->max_level || !T.inside(lev[e], ix[e], iy[e])) return -3;
    base[e] = Cell{lev[e], ix[e], iy[e]};
    Node nd{{h[e], hu[e], hv[e]}, el_error ? el_error[e] : 0., NOTRECADAPTED};
    if (!T.leaf.emplace(ckey(lev[e], ix[e], iy[e]), nd).second) return -3;
  }
  // the leaves must tile the base grid and be 2:1 balanced
  {
    double area = 0;
    for (const Cell& c : base) area += 1.0 / (double)(1ull << (2 * c.lev));
    if (area != (double)nx0 * ny0) return -3;
    for (const Cell& c : base)
      for (int s = 0; s < 4; s++) {
        int nl[2];
        int64_t ni[2], nj[2];
        if (T.side_nbrs(c.lev, c.i, c.j, s, nl, ni, nj) < 0) return -3;
      }
  }
  const double tiny = prm->tiny;
  const double contours[4] = {tiny, 5 * tiny, 15 * tiny, 40 * tiny};  // hadpt.C:46-48
  // element_weight (element_weight.C:36-79): mean error of the leaves with
  // hu > tiny
  double geo_target = 1.0;
  if (el_error) {
    double s = 0, w = 0;
    for (int e = 0; e < n; e++)
      if (hu[e] > tiny) {
        s += el_error[e];
        w += 1;
      }
    if (w > 0) geo_target = s / w;
  }
  // refinement along the pile contours (and where the error is large)
  std::vector<Cell> cur, todo;
  ordered_leaves(T, base, cur);
  for (const Cell& c : cur) {
    const Node* nd = T.at(c.lev, c.i, c.j);
    if (!(nd->flag > 0 && nd->flag < NEWSON) || c.lev >= T.max_level) continue;
    bool ref = el_error && nd->err > geo_target;
    for (int k = 0; k < 4 && !ref; k++) ref = pile_boundary(T, c.lev, c.i, c.j, contours[k]) > 0;
    if (ref) todo.push_back(c);
  }
  for (const Cell& c : todo) T.refine(c.lev, c.i, c.j);
  // buffer layers around the contours
  if (prm->num_buffer_layer >= 1) {
    ordered_leaves(T, base, cur);
    for (const Cell& c : cur) {
      if (!T.at(c.lev, c.i, c.j)) continue;  // refined by a cascade meanwhile
      bool ref = false;
      for (int k = 0; k < 4 && !ref; k++) ref = first_buffer(T, c.lev, c.i, c.j, contours[k]) == 1;
      if (ref) T.refine(c.lev, c.i, c.j);
    }
    ordered_leaves(T, base, cur);
    todo.clear();
    for (const Cell& c : cur) {
      bool buf = false;
      for (int k = 0; k < 4 && !buf; k++) buf = first_buffer(T, c.lev, c.i, c.j, contours[k]) > 0;
      if (buf) todo.push_back(c);
    }
    for (const Cell& c : todo) T.leaf[ckey(c.lev, c.i, c.j)].flag = BUFFER;
    for (int layer = 2; layer <= prm->num_buffer_layer; layer++) {
      ordered_leaves(T, base, cur);
      for (const Cell& c : cur)
        if (T.at(c.lev, c.i, c.j) && next_buffer(T, c.lev, c.i, c.j) == 1) T.refine(c.lev, c.i, c.j);
      ordered_leaves(T, base, cur);
      todo.clear();
      for (const Cell& c : cur)
        if (next_buffer(T, c.lev, c.i, c.j) > 0) todo.push_back(c);
      for (const Cell& c : todo) T.leaf[ckey(c.lev, c.i, c.j)].flag = BUFFER;
    }
  }
  // unrefinement: four untouched brothers none of which has a finer
  // neighbour become their father (find_brothers / check_unrefinement,
  // unrefine.C:211-250, 336-352); decided on the mesh before any merge
  ordered_leaves(T, base, cur);
  todo.clear();
  for (const Cell& c : cur) {
    if (c.lev < 1 || (c.i & 1) || (c.j & 1)) continue;  // son 0 checks its family
    bool ok = true;
    for (int k = 0; k < 4 && ok; k++) {
      const int64_t bi = c.i + (k == 1 || k == 2), bj = c.j + (k >= 2);
      const Node* b = T.at(c.lev, bi, bj);
      if (!b || b->flag != NOTRECADAPTED) {
        ok = false;
        break;
      }
      for (int s = 0; s < 4 && ok; s++) {
        int nl[2];
        int64_t ni[2], nj[2];
        if (T.side_nbrs(c.lev, bi, bj, s, nl, ni, nj) == 2) ok = false;
      }
    }
    if (ok) todo.push_back(c);
  }
  for (const Cell& c : todo) {
    static const int SI[4] = {0, 1, 1, 0}, SJ[4] = {0, 0, 1, 1};
    Node f{{0., 0., 0.}, 0., NEWFATHER};
    for (int k = 0; k < 4; k++) {
      const uint64_t kk = ckey(c.lev, c.i + SI[k], c.j + SJ[k]);
      const Node& s = T.leaf[kk];
      for (int v = 0; v < 3; v++) f.U[v] += s.U[v] * 0.25;
      f.err += s.err;
      T.leaf.erase(kk);
    }
    T.leaf[ckey(c.lev - 1, c.i >> 1, c.j >> 1)] = f;
    T.ops.push_back(tg_adapt_op{1, c.lev - 1, (int32_t)(c.i >> 1), (int32_t)(c.j >> 1)});
  }
  // the new leaves, in order: a new father where its son 0 was (cur is the
  // order before the merges)
  std::vector<Cell> out;
  for (const Cell& c : cur) {
    if (T.at(c.lev, c.i, c.j)) out.push_back(c);
    else if (!(c.i & 1) && !(c.j & 1) && T.at(c.lev - 1, c.i >> 1, c.j >> 1))
      out.push_back(Cell{c.lev - 1, c.i >> 1, c.j >> 1});
  }
  *n_out = (int)out.size();
  *nops = (int)T.ops.size();
  if ((int)out.size() > cap || (int)T.ops.size() > op_cap) return -2;
  if (!lev_out || !ix_out || !iy_out || !h_out || !hu_out || !hv_out || (!ops && *nops > 0))
    return -1;
  for (size_t e = 0; e < out.size(); e++) {
    const Node* nd = T.at(out[e].lev, out[e].i, out[e].j);
    lev_out[e] = out[e].lev; ix_out[e] = (int32_t)out[e].i; iy_out[e] = (int32_t)out[e].j;
    h_out[e] = nd->U[0]; hu_out[e] = nd->U[1]; hv_out[e] = nd->U[2];
  }
  if (*nops > 0) std::memcpy(ops, T.ops.data(), sizeof(tg_adapt_op) * T.ops.size());
  return 0;
}

extern "C" int tg_host_adapt_adjoint(int nops, const tg_adapt_op* ops, int n_new,
                                     const int32_t* lev_new, const int32_t* ix_new,
                                     const int32_t* iy_new, const double* lam_new, int n_old,
                                     const int32_t* lev_old, const int32_t* ix_old,
                                     const int32_t* iy_old, double* lam_old) {
  if (nops < 0 || (nops > 0 && !ops) || n_new < 0 || n_old < 0 ||
      (n_new > 0 && (!lev_new || !ix_new || !iy_new || !lam_new)) ||
      (n_old > 0 && (!lev_old || !ix_old || !iy_old || !lam_old)))
    return -1;
  struct L3 {
    double v[3];
  };
  std::unordered_map<uint64_t, L3> lam;
  for (int e = 0; e < n_new; e++) {
    L3 x{{lam_new[e], lam_new[(size_t)n_new + e], lam_new[2 * (size_t)n_new + e]}};
    if (!lam.emplace(ckey(lev_new[e], ix_new[e], iy_new[e]), x).second) return -3;
  }
  static const int SI[4] = {0, 1, 1, 0}, SJ[4] = {0, 0, 1, 1};
  for (int q = nops - 1; q >= 0; q--) {
    const tg_adapt_op& o = ops[q];
    const int64_t i = o.ix, j = o.iy;
    if (o.kind == 0) {  // a split: the father's adjoint is the sum of its sons'
      L3 f{{0., 0., 0.}};
      for (int k = 0; k < 4; k++) {
        auto it = lam.find(ckey(o.lev + 1, 2 * i + SI[k], 2 * j + SJ[k]));
        if (it == lam.end()) return -3;
        for (int v = 0; v < 3; v++) f.v[v] += it->second.v[v];
        lam.erase(it);
      }
      if (!lam.emplace(ckey(o.lev, i, j), f).second) return -3;
    } else if (o.kind == 1) {  // a merge: each son takes 0.25 of the father's
      auto it = lam.find(ckey(o.lev, i, j));
      if (it == lam.end()) return -3;
      const L3 f = it->second;
      lam.erase(it);
      for (int k = 0; k < 4; k++) {
        L3 s;
        for (int v = 0; v < 3; v++) s.v[v] = 0.25 * f.v[v];
        if (!lam.emplace(ckey(o.lev + 1, 2 * i + SI[k], 2 * j + SJ[k]), s).second) return -3;
      }
    } else {
      return -1;
    }
  }
  if ((int)lam.size() != n_old) return -3;
  for (int e = 0; e < n_old; e++) {
    auto it = lam.find(ckey(lev_old[e], ix_old[e], iy_old[e]));
    if (it == lam.end()) return -3;
    for (int v = 0; v < 3; v++) lam_old[(size_t)v * n_old + e] = it->second.v[v];
  }
  return 0;
}

extern "C" int tg_host_adapt_forward(int nops, const tg_adapt_op* ops, int nf, int n_old,
                                     const int32_t* lev_old, const int32_t* ix_old,
                                     const int32_t* iy_old, const double* f_old, int n_new,
                                     const int32_t* lev_new, const int32_t* ix_new,
                                     const int32_t* iy_new, double* f_new) {
  if (nops < 0 || (nops > 0 && !ops) || nf < 1 || n_old < 0 || n_new < 0 ||
      (n_old > 0 && (!lev_old || !ix_old || !iy_old || !f_old)) ||
      (n_new > 0 && (!lev_new || !ix_new || !iy_new || !f_new)))
    return -1;
  std::unordered_map<uint64_t, std::vector<double>> val;
  for (int e = 0; e < n_old; e++) {
    std::vector<double> x(nf);
    for (int v = 0; v < nf; v++) x[v] = f_old[(size_t)v * n_old + e];
    if (!val.emplace(ckey(lev_old[e], ix_old[e], iy_old[e]), std::move(x)).second) return -3;
  }
  static const int SI[4] = {0, 1, 1, 0}, SJ[4] = {0, 0, 1, 1};
  for (int q = 0; q < nops; q++) {
    const tg_adapt_op& o = ops[q];
    const int64_t i = o.ix, j = o.iy;
    if (o.kind == 0) {  // split: every son copies the father (x 1.0)
      auto it = val.find(ckey(o.lev, i, j));
      if (it == val.end()) return -3;
      const std::vector<double> f = it->second;
      val.erase(it);
      for (int k = 0; k < 4; k++) {
        std::vector<double> s(nf);
        for (int v = 0; v < nf; v++) s[v] = f[v] * 1.0;
        if (!val.emplace(ckey(o.lev + 1, 2 * i + SI[k], 2 * j + SJ[k]), std::move(s)).second)
          return -3;
      }
    } else if (o.kind == 1) {  // merge: the father sums 0.25 x its sons in order
      std::vector<double> f(nf, 0.);
      for (int k = 0; k < 4; k++) {
        auto it = val.find(ckey(o.lev + 1, 2 * i + SI[k], 2 * j + SJ[k]));
        if (it == val.end()) return -3;
        for (int v = 0; v < nf; v++) f[v] += it->second[v] * 0.25;
        val.erase(it);
      }
      if (!val.emplace(ckey(o.lev, i, j), std::move(f)).second) return -3;
    } else {
      return -1;
    }
  }
  if ((int)val.size() != n_new) return -3;
  for (int e = 0; e < n_new; e++) {
    auto it = val.find(ckey(lev_new[e], ix_new[e], iy_new[e]));
    if (it == val.end()) return -3;
    for (int v = 0; v < nf; v++) f_new[(size_t)v * n_new + e] = it->second[v];
  }
  return 0;
}
