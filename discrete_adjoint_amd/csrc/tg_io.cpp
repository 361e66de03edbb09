// The reference's run inputs (Read_data, datread.C:36-451): simulation.data
// (sources and piles, cells across the smallest pile, time limits, flags,
// GIS names, test points, discharge planes), scale.data (length, height and
// gravity scales) and frict.data (materials with their friction angles),
// read token by token in the reference's order and non-dimensionalised as it
// does.  Host code: it fills what tg_host_piles, tg_params and the adaptive
// time loop take.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/tg_host.h"

namespace {

constexpr double kPi = 3.14159265358979323846;  // PI of the reference's constant.h

template <class T>
bool next(std::ifstream& in, T* v) {
  return static_cast<bool>(in >> *v);
}

void copy_name(char* dst, size_t cap, const std::string& s) {
  std::strncpy(dst, s.c_str(), cap - 1);
  dst[cap - 1] = 0;
}

}  // namespace

extern "C" int tg_host_read_run_inputs(const char* dir, int pile_cap, tg_pile* piles,
                                       tg_run_inputs* out) {
  if (!dir || !out || pile_cap < 0 || (pile_cap > 0 && !piles)) return -1;
  std::memset(out, 0, sizeof(*out));
  const std::string d = std::string(dir) + "/";
  std::ifstream sim(d + "simulation.data");
  if (!sim) return -4;
  if (!next(sim, &out->srctype)) return -5;
  if (out->srctype & 0x1 && !next(sim, &out->npiles)) return -5;
  if (out->srctype & 0x2 && !next(sim, &out->nsources)) return -5;
  if (out->npiles < 0 || out->nsources < 0) return -5;
  std::vector<tg_pile> pv(out->npiles);
  for (int p = 0; p < out->npiles; p++) {
    double rot, speed, dir_deg;
    tg_pile& q = pv[p];
    if (!(next(sim, &q.height) && next(sim, &q.xc) && next(sim, &q.yc) &&
          next(sim, &q.major) && next(sim, &q.minor) && next(sim, &rot) && next(sim, &speed) &&
          next(sim, &dir_deg)))
      return -5;
    q.cosrot = std::cos(rot * kPi / 180.0);
    q.sinrot = std::sin(rot * kPi / 180.0);
    q.rot_deg = rot;
    const double a = dir_deg * kPi / 180.0;
    q.vx = speed * std::cos(a);
    q.vy = speed * std::sin(a);
  }
  // flux sources: influx, start, end, centre, radii, rotation, speed, angle
  for (int s = 0; s < out->nsources; s++)
    for (int k = 0; k < 10; k++) {
      double skip;
      if (!next(sim, &skip)) return -5;
    }
  // scale.data (absent: no scaling)
  out->length_scale = out->height_scale = out->gravity_scale = 1.0;
  {
    std::ifstream sc(d + "scale.data");
    if (sc) {
      if (!(next(sc, &out->length_scale) && next(sc, &out->height_scale) &&
            next(sc, &out->gravity_scale)))
        return -5;
    }
  }
  out->epsilon = out->height_scale / out->length_scale;
  out->time_scale = std::sqrt(out->length_scale / out->gravity_scale);
  const double vscale = std::sqrt(out->length_scale * out->gravity_scale);
  double smallest = HUGE_VAL;
  for (tg_pile& q : pv) {
    q.height /= out->height_scale;
    q.xc /= out->length_scale;
    q.yc /= out->length_scale;
    q.major /= out->length_scale;
    q.minor /= out->length_scale;
    q.vx /= vscale;
    q.vy /= vscale;
    if (smallest > q.major) smallest = q.major;
    if (smallest > q.minor) smallest = q.minor;
  }
  out->smallest_axis = 2.0 * smallest;
  if (!next(sim, &out->cells_across)) return -5;
  // time limits (TimeProps::inittime, properties.h:375-394)
  if (!(next(sim, &out->maxiter) && next(sim, &out->maxtime) && next(sim, &out->timeoutput) &&
        next(sim, &out->timesave)))
    return -5;
  out->nd_maxtime = out->maxtime / out->time_scale;
  out->nd_nextoutput = out->timeoutput / out->time_scale;
  out->nd_nextsave = out->timesave / out->time_scale;
  if (!(next(sim, &out->adaptflag) && next(sim, &out->vizflag) && next(sim, &out->order)))
    return -5;
  std::string gmain, gsub, gmapset, gmap;
  if (!(next(sim, &gmain) && next(sim, &gsub) && next(sim, &gmapset) && next(sim, &gmap) &&
        next(sim, &out->extramaps)))
    return -5;
  copy_name(out->gis_main, sizeof out->gis_main, gmain);
  copy_name(out->gis_sub, sizeof out->gis_sub, gsub);
  copy_name(out->gis_mapset, sizeof out->gis_mapset, gmapset);
  copy_name(out->gis_map, sizeof out->gis_map, gmap);
  // test point (datread.C:302-330): -1 = default heights
  const double max_negligible = out->height_scale / 10000.0;
  if (!next(sim, &out->edge_height)) return -5;
  if (out->edge_height == -1.0) out->edge_height = max_negligible * 10.0;
  else if (out->edge_height <= 0.0) return -5;
  out->edge_height /= out->height_scale;
  if (!next(sim, &out->test_height)) return -5;
  if (out->test_height != -2) {
    if (out->test_height == -1) out->test_height = max_negligible * 10.0;
    out->test_height /= out->height_scale;
    if (!(next(sim, &out->test_x) && next(sim, &out->test_y))) return -5;
    out->test_x /= out->length_scale;
    out->test_y /= out->length_scale;
  } else {
    std::string g1, g2;
    if (!(next(sim, &g1) && next(sim, &g2))) return -5;
    out->test_height = out->test_x = out->test_y = HUGE_VAL;
  }
  // discharge planes: four coordinates each
  if (!next(sim, &out->nplanes) || out->nplanes < 0) return -5;
  for (int p = 0; p < out->nplanes; p++)
    for (int k = 0; k < 4; k++) {
      double skip;
      if (!next(sim, &skip)) return -5;
    }
  // frict.data (datread.C:379-403): a count, then per material a name line
  // and "internal bed" angles in degrees; the internal friction is the first
  // material's
  {
    std::ifstream fr(d + "frict.data");
    if (!fr) return -4;
    if (!next(fr, &out->nmat) || out->nmat < 1 || out->nmat > TG_HOST_MAX_MATERIALS) return -5;
    std::string line;
    std::getline(fr, line);  // rest of the count line
    for (int m = 1; m <= out->nmat; m++) {
      if (!std::getline(fr, line)) return -5;
      copy_name(out->matnames[m - 1], sizeof out->matnames[0], line);
      double intf, bedf;
      if (!(next(fr, &intf) && next(fr, &bedf))) return -5;
      std::getline(fr, line);
      out->bedfrict[m - 1] = bedf * kPi / 180.0;
      if (m == 1) out->intfrict = intf * kPi / 180.0;
    }
  }
  if (out->npiles > pile_cap) return -2;
  for (int p = 0; p < out->npiles; p++) piles[p] = pv[p];
  return 0;
}
