// The reference's grid file funky%04d.inp (Read_grid, datread.C:460-660),
// as its preprocessor writes it (preprocess.C:560-660, node.C:82-93,
// element2.C:233-283): binary in the writer's byte order.
//   int version: 20061109 (coordinates and ranges as float) or 20061110
//                (as double); int node count; key range [2], x range [2],
//                y range [2]
//   per node:    unsigned key[2]; x, y
//   int element count
//   per element: unsigned key[2] of each of its 9 nodes (node 8: the
//                bubble node, the element's own key); per side 0..3: int
//                neighbour process (-1: domain boundary) then, unless -1,
//                unsigned key[2] of the neighbour; per side: int essential
//                boundary flag (-1: none) then, unless -1, float x, y value;
//                the same for the natural boundary condition; int elm_loc[2]
//                (the element's i, j in the base grid); unsigned opposite
//                brother key[2]; int material
//   (then the material names, which Read_grid does not read)
// Host code: lengths are divided by the length scale, as Read_grid does.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/tg_host.h"

namespace {

struct Reader {
  FILE* fp;
  bool ok = true;
  template <class T>
  T get() {
    T v{};
    if (ok && std::fread(&v, sizeof(T), 1, fp) != 1) ok = false;
    return v;
  }
  double real(bool floats) { return floats ? (double)get<float>() : get<double>(); }
};

}  // namespace

extern "C" int tg_host_read_funky(const char* path, double length_scale, tg_funky_header* hdr,
                                  int node_cap, uint32_t* node_keys, double* node_xy,
                                  int elem_cap, tg_funky_elem* elems) {
  if (!path || !hdr || !(length_scale > 0.) || node_cap < 0 || elem_cap < 0 ||
      (node_cap > 0 && (!node_keys || !node_xy)) || (elem_cap > 0 && !elems))
    return -1;
  std::memset(hdr, 0, sizeof(*hdr));
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return -4;
  Reader r{fp};
  int rc = 0;
  hdr->version = r.get<int32_t>();
  const bool floats = hdr->version == 20061109;
  if (!r.ok || (!floats && hdr->version != 20061110)) rc = -5;
  if (!rc) {
    hdr->nnodes = r.get<int32_t>();
    for (int i = 0; i < 2; i++) hdr->keyrange[i] = r.real(floats);
    for (int i = 0; i < 2; i++) hdr->xrange[i] = r.real(floats) / length_scale;
    for (int i = 0; i < 2; i++) hdr->yrange[i] = r.real(floats) / length_scale;
    if (!r.ok || hdr->nnodes < 0) rc = -5;
  }
  const bool fill_nodes = !rc && hdr->nnodes <= node_cap;
  for (int n = 0; !rc && n < hdr->nnodes; n++) {
    const uint32_t k0 = r.get<uint32_t>(), k1 = r.get<uint32_t>();
    const double x = r.real(floats) / length_scale, y = r.real(floats) / length_scale;
    if (!r.ok) rc = -5;
    if (fill_nodes) {
      node_keys[2 * n] = k0; node_keys[2 * n + 1] = k1;
      node_xy[2 * n] = x; node_xy[2 * n + 1] = y;
    }
  }
  if (!rc) {
    hdr->nelems = r.get<int32_t>();
    if (!r.ok || hdr->nelems < 0) rc = -5;
  }
  const bool fill_elems = !rc && hdr->nelems <= elem_cap;
  for (int e = 0; !rc && e < hdr->nelems; e++) {
    tg_funky_elem el;
    std::memset(&el, 0, sizeof(el));
    for (int j = 0; j < 9; j++)
      for (int k = 0; k < 2; k++) el.node_key[j][k] = r.get<uint32_t>();
    for (int j = 0; j < 4; j++) {
      el.neigh_proc[j] = r.get<int32_t>();
      if (el.neigh_proc[j] != -1)
        for (int k = 0; k < 2; k++) el.neigh_key[j][k] = r.get<uint32_t>();
    }
    // Read_grid's BC types: 1 essential, 2 natural, 3 both (0: none)
    for (int kind = 0; kind < 2; kind++)
      for (int j = 0; j < 4; j++) {
        if (r.get<int32_t>() == -1) continue;
        el.bc_type[j] = kind == 0 ? 1 : (el.bc_type[j] == 0 ? 2 : 3);
        for (int k = 0; k < 2; k++) el.bc_value[j][kind][k] = r.get<float>();
      }
    for (int k = 0; k < 2; k++) el.elm_loc[k] = r.get<int32_t>();
    for (int k = 0; k < 2; k++) el.opposite_brother[k] = r.get<uint32_t>();
    el.material = r.get<int32_t>();
    if (!r.ok) rc = -5;
    if (fill_elems) elems[e] = el;
  }
  std::fclose(fp);
  if (rc) return rc;
  return (fill_nodes && fill_elems) ? 0 : -2;
}
