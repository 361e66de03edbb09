// GRASS raster input (the reference's GIS reader: GisRasterHdr,
// GisBinFile, GisApi.C:153-285, 1455-1480): the cellhd header of an
// elevation map and its fcell data -- big-endian float32 rows, raw or
// row-compressed (row offsets of 4 or 8 bytes, each row flagged '1' and
// zlib-deflated, or stored as is).  The result is the tg_dem that
// tg_host_static_coefs and tg_host_leaf_static_coefs take.
#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tg_host.h"

namespace {

struct Header {
  int proj = 0, zone = 0, cols = 0, rows = 0, format = 0, compressed = 0;
  double north = 0, south = 0, east = 0, west = 0, ewres = 0, nsres = 0;
};

// GisRasterHdr: twelve "key: value" lines in a fixed order, read by position
// after the first key is checked to be "proj"
bool read_header(const std::string& path, Header* h) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  std::vector<std::string> vals;
  char line[512];
  bool first = true, ok = true;
  while (std::fgets(line, sizeof line, f) && vals.size() < 12) {
    char* colon = std::strchr(line, ':');
    if (!colon) continue;
    if (first) {
      std::string key(line, colon - line);
      ok = key == "proj";
      first = false;
      if (!ok) break;
    }
    vals.emplace_back(colon + 1);
  }
  std::fclose(f);
  if (!ok || vals.size() < 12) return false;
  auto I = [&](int k) { return (int)std::strtol(vals[k].c_str(), nullptr, 10); };
  auto D = [&](int k) { return std::strtod(vals[k].c_str(), nullptr); };
  h->proj = I(0); h->zone = I(1);
  h->north = D(2); h->south = D(3); h->east = D(4); h->west = D(5);
  h->cols = I(6); h->rows = I(7);
  h->ewres = D(8); h->nsres = D(9);
  h->format = I(10); h->compressed = I(11);
  return true;
}

float be_float(const unsigned char* p) {
  uint32_t u = ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
  float v;
  std::memcpy(&v, &u, 4);
  return v;
}

bool be_offset(const unsigned char* p, int n, int64_t* out) {
  uint64_t v = 0;
  for (int i = 0; i < n; i++) v = (v << 8) | p[i];
  if (v > (uint64_t)INT64_MAX) return false;
  *out = (int64_t)v;
  return true;
}

// GisBinFile::readRow / readCompressdRow for float rows
int read_rows(const std::string& path, const Header& h, float* elev) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return -4;
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<unsigned char> buf(size > 0 ? (size_t)size : 0);
  const bool got = size > 0 && std::fread(buf.data(), 1, buf.size(), f) == buf.size();
  std::fclose(f);
  if (!got) return -5;
  const size_t rowbytes = (size_t)h.cols * 4;
  std::vector<unsigned char> row(rowbytes);
  for (int r = 0; r < h.rows; r++) {
    const unsigned char* src;
    if (!h.compressed) {
      if ((size_t)(r + 1) * rowbytes > buf.size()) return -5;
      src = buf.data() + (size_t)r * rowbytes;
    } else {
      const int nb = buf[0];
      if (nb != 4 && nb != 8) return -5;
      const size_t at = 1 + (size_t)r * nb;
      if (at + 2 * (size_t)nb > buf.size()) return -5;
      int64_t p0, p1;
      if (!be_offset(buf.data() + at, nb, &p0) || !be_offset(buf.data() + at + nb, nb, &p1))
        return -5;
      if (p0 < 0 || p1 <= p0 || (size_t)p1 > buf.size()) return -5;
      const unsigned char flag = buf[(size_t)p0];
      const size_t len = (size_t)(p1 - p0 - 1);
      if (flag == '1') {  // zlib-deflated row (GisUncompress)
        uLongf out = (uLongf)rowbytes;
        if (uncompress(row.data(), &out, buf.data() + p0 + 1, (uLong)len) != Z_OK ||
            out != rowbytes)
          return -5;
      } else {  // stored row
        if (len < rowbytes) return -5;
        std::memcpy(row.data(), buf.data() + p0 + 1, rowbytes);
      }
      src = row.data();
    }
    for (int c = 0; c < h.cols; c++) elev[(size_t)r * h.cols + c] = be_float(src + 4 * c);
  }
  return 0;
}

}  // namespace

extern "C" int tg_host_read_grass(const char* gisdbase, const char* location, const char* mapset,
                                  const char* name, long long capacity, tg_dem* dem,
                                  float* elev) {
  if (!gisdbase || !location || !mapset || !name || !dem) return -1;
  const std::string dir = std::string(gisdbase) + "/" + location + "/" + mapset + "/";
  Header h;
  if (!read_header(dir + "cellhd/" + name, &h)) return -4;
  if (h.rows < 1 || h.cols < 1) return -4;
  // set_from_header (GisApi.C:1455-1480): one resolution, edges from the
  // west / south corner
  if (std::fabs(h.ewres - h.nsres) >= h.ewres / 10.0) return -4;
  dem->nrows = h.rows;
  dem->ncols = h.cols;
  dem->res = h.ewres;
  dem->xmin = h.west;
  dem->ymax = h.south + h.rows * h.nsres;
  dem->elev = elev;
  if ((long long)h.rows * h.cols > capacity || !elev) return -2;
  return read_rows(dir + "fcell/" + name, h, elev);
}
