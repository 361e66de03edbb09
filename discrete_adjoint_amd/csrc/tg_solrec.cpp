// Trajectory spill files in the reference's SolRec layout, one gzip stream
// per step (SolRec::wrtie_sol_to_disk dualmesh.C:87-142, read_sol_from_disk
// dualmesh.C:234-297): int32 nzero, int32 nnz, nzero element keys (2 x
// uint32), then nnz records of key (2 x uint32) and h, hu, hv (FP64), in the
// host's byte order.  An element whose h > 0 is false is a zero entry
// (SolRec::record_solution dualmesh.C:53-70) and reads back as (0, 0, 0).
#include "../../include/tg_host.h"

#include <zlib.h>

#include <cstdint>
#include <unordered_map>
#include <vector>

namespace {

struct GzFile {
  gzFile f;
  explicit GzFile(gzFile g) : f(g) {}
  ~GzFile() {
    if (f) gzclose(f);
  }
};

bool put(gzFile f, const void* p, unsigned bytes) {
  return gzwrite(f, p, bytes) == (int)bytes;
}
bool get(gzFile f, void* p, unsigned bytes) {
  return gzread(f, p, bytes) == (int)bytes;
}

uint64_t key64(const uint32_t* k) { return ((uint64_t)k[0] << 32) | k[1]; }

}  // namespace

extern "C" int tg_host_write_solution(const char* path, int64_t n, const uint32_t* keys,
                                      const int32_t* order, const double* h, const double* hu,
                                      const double* hv) {
  if (!path || n < 0 || n > INT32_MAX || (n > 0 && (!keys || !h || !hu || !hv))) return -1;
  std::vector<int64_t> zero, nonzero;
  for (int64_t q = 0; q < n; q++) {
    const int64_t e = order ? order[q] : q;
    if (e < 0 || e >= n) return -1;
    if (h[e] > 0.) nonzero.push_back(e);
    else zero.push_back(e);
  }
  GzFile g(gzopen(path, "wb"));
  if (!g.f) return -2;
  const int32_t nz = (int32_t)zero.size(), nnz = (int32_t)nonzero.size();
  if (!put(g.f, &nz, 4) || !put(g.f, &nnz, 4)) return -2;
  for (int64_t e : zero)
    if (!put(g.f, keys + 2 * e, 8)) return -2;
  for (int64_t e : nonzero) {
    const double s[3] = {h[e], hu[e], hv[e]};
    if (!put(g.f, keys + 2 * e, 8) || !put(g.f, s, 24)) return -2;
  }
  const int rc = gzclose(g.f);
  g.f = nullptr;
  return rc == Z_OK ? 0 : -2;
}

extern "C" int tg_host_read_solution(const char* path, int64_t n, const uint32_t* keys,
                                     double* h, double* hu, double* hv, int64_t* found,
                                     int64_t* unknown) {
  if (!path || n < 0 || (n > 0 && (!keys || !h || !hu || !hv))) return -1;
  std::unordered_map<uint64_t, int64_t> at;
  at.reserve((size_t)n * 2);
  for (int64_t e = 0; e < n; e++) at.emplace(key64(keys + 2 * e), e);
  GzFile g(gzopen(path, "rb"));
  if (!g.f) return -2;
  int32_t nz = 0, nnz = 0;
  if (!get(g.f, &nz, 4) || !get(g.f, &nnz, 4) || nz < 0 || nnz < 0) return -3;
  int64_t hit = 0, miss = 0;
  for (int64_t i = 0; i < (int64_t)nz + nnz; i++) {
    uint32_t k[2];
    double s[3] = {0., 0., 0.};
    if (!get(g.f, k, 8)) return -3;
    if (i >= nz && !get(g.f, s, 24)) return -3;
    const auto it = at.find(key64(k));
    if (it == at.end()) {
      miss++;
      continue;
    }
    h[it->second] = s[0];
    hu[it->second] = s[1];
    hv[it->second] = s[2];
    hit++;
  }
  char extra;
  if (gzread(g.f, &extra, 1) != 0) return -3;  // trailing bytes: not this format
  if (found) *found = hit;
  if (unknown) *unknown = miss;
  return 0;
}
