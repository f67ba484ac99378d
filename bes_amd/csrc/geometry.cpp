// Host-side mesh topology and metrics (see geometry.hpp).
//
// The face numbering, normal orientation and metric formulas reproduce the
// reference reader and element classes, quirks included, so that the device
// kernels see exactly the reference's geometry:
//   face discovery and orientation   Mesh.cpp:236-534, MeshCommon.cpp:60-96
//   face centre / area vector        MeshGmsh.cpp:26-68
//   cell volume / centroid           MeshGmsh.cpp:76-145
//   centroid->face vectors, IDW      Element.cpp:111-140
//   adjacency                        Mesh.cpp:601-634
//   projected lengths                TimeStepCalculatorLocal.cpp:51-65
// Instead of the reference's string-keyed map, faces are matched by sorting
// packed vertex keys, then numbered in the reference's discovery order.
#include "geometry.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace bes {
namespace {

const int kTetFace[4][3] = {{0, 2, 1}, {1, 2, 3}, {0, 3, 2}, {0, 1, 3}};

struct Key {
  uint64_t hi, lo;
  bool operator<(const Key &o) const { return hi < o.hi || (hi == o.hi && lo < o.lo); }
  bool operator==(const Key &o) const { return hi == o.hi && lo == o.lo; }
};

Key make_key(const int64_t *v, int nv) {
  int64_t s[3] = {v[0], v[1], nv > 2 ? v[2] : -1};
  if (nv == 2) {
    if (s[0] > s[1]) std::swap(s[0], s[1]);
    return Key{(uint64_t)s[0], (uint64_t)s[1]};
  }
  std::sort(s, s + 3);
  return Key{(uint64_t)s[0], ((uint64_t)s[1] << 32) | (uint64_t)s[2]};
}

// (c - b) x (a - b)
void cross3(double *r, const double *a, const double *b, const double *c) {
  r[0] = (c[1] - b[1]) * (a[2] - b[2]) - (a[1] - b[1]) * (c[2] - b[2]);
  r[1] = (a[0] - b[0]) * (c[2] - b[2]) - (c[0] - b[0]) * (a[2] - b[2]);
  r[2] = (c[0] - b[0]) * (a[1] - b[1]) - (a[0] - b[0]) * (c[1] - b[1]);
}

double norm_pow(const double *v, int nd) {
  double s = 0;
  for (int i = 0; i < nd; i++) s += std::pow(v[i], 2.0);
  return std::sqrt(s);
}

double dist_pow(const double *a, const double *b, int nd) {
  double s = 0;
  for (int i = 0; i < nd; i++) s += std::pow(a[i] - b[i], 2.0);
  return std::sqrt(s);
}

// face metrics from an ordered vertex list
void face_metrics(const double *X, int nd, const int64_t *fv, double *S, double *nrm,
                  double *area, double *ctr) {
  double s[3] = {0, 0, 0};
  if (nd == 2) {
    const double *a = X + fv[0] * 2, *b = X + fv[1] * 2;
    for (int i = 0; i < 2; i++) ctr[i] = 0.5 * (a[i] + b[i]);
    s[0] = b[1] - a[1];
    s[1] = a[0] - b[0];
  } else {
    const double *a = X + fv[0] * 3, *b = X + fv[1] * 3, *c = X + fv[2] * 3;
    for (int i = 0; i < 3; i++) ctr[i] = (a[i] + b[i] + c[i]) / 3;
    cross3(s, a, b, c);
    for (int i = 0; i < 3; i++) s[i] /= 2.0;
  }
  double ar = norm_pow(s, nd);
  *area = ar;
  for (int i = 0; i < nd; i++) {
    S[i] = s[i];
    nrm[i] = s[i] / ar;
  }
}

void local_face(const int64_t *cv, int nd, int m, int64_t *fv) {
  if (nd == 2) {
    fv[0] = cv[m];
    fv[1] = cv[(m + 1) % 3];
  } else {
    for (int q = 0; q < 3; q++) fv[q] = cv[kTetFace[m][q]];
  }
}

}  // namespace

Geometry build_geometry(const MeshInput &in, int mode) {
  const int nd = in.ndim;
  if (nd != 2 && nd != 3) throw std::runtime_error("mesh dimension must be 2 or 3");
  const int k = nd + 1;
  const int64_t ne = in.ncell(), nb = in.nbface(), nn = in.nnode();
  if (ne <= 0) throw std::runtime_error("mesh has no cells");
  const double *X = in.coords.data();
  for (int64_t q = 0; q < ne * k; q++)
    if (in.cells[q] < 0 || in.cells[q] >= nn) throw std::runtime_error("cell node id out of range");
  for (int64_t q = 0; q < nb * nd; q++)
    if (in.bfaces[q] < 0 || in.bfaces[q] >= nn) throw std::runtime_error("face node id out of range");

  Geometry g;
  g.nd = nd; g.k = k; g.mode = mode; g.ncell = ne;

  // ---- match every (cell, local face) slot with its boundary face or twin ----
  const int64_t ns = ne * k;
  struct Slot { Key key; int64_t idx; };
  std::vector<Slot> slots(ns);
  for (int64_t e = 0; e < ne; e++)
    for (int m = 0; m < k; m++) {
      int64_t fv[3];
      local_face(&in.cells[e * k], nd, m, fv);
      slots[e * k + m] = Slot{make_key(fv, nd), e * k + m};
    }
  std::sort(slots.begin(), slots.end(), [](const Slot &a, const Slot &b) {
    return a.key < b.key || (a.key == b.key && a.idx < b.idx);
  });
  std::vector<std::pair<Key, int64_t>> bkeys(nb);
  for (int64_t b = 0; b < nb; b++) bkeys[b] = {make_key(&in.bfaces[b * nd], nd), b};
  std::sort(bkeys.begin(), bkeys.end(), [](const std::pair<Key, int64_t> &a, const std::pair<Key, int64_t> &b) {
    return a.first < b.first || (a.first == b.first && a.second < b.second);
  });

  g.cface.assign(ns, -1);
  // internal faces: (first slot, second slot)
  std::vector<std::pair<int64_t, int64_t>> internal;
  internal.reserve(ns / 2 + 1);
  std::vector<std::vector<int64_t>> bslots(nb);   // slots touching each boundary face
  size_t bi = 0;
  for (int64_t s0 = 0; s0 < ns;) {
    int64_t s1 = s0 + 1;
    while (s1 < ns && slots[s1].key == slots[s0].key) s1++;
    while (bi < bkeys.size() && bkeys[bi].first < slots[s0].key) bi++;
    // the reference keeps only the last boundary element with a given key
    int64_t bmatch = -1;
    for (size_t q = bi; q < bkeys.size() && bkeys[q].first == slots[s0].key; q++) bmatch = bkeys[q].second;
    if (bmatch >= 0) {
      for (int64_t s = s0; s < s1; s++) {
        bslots[bmatch].push_back(slots[s].idx);
        g.cface[slots[s].idx] = bmatch;
      }
    } else {
      int64_t a = slots[s0].idx, b = (s1 - s0 > 1) ? slots[s1 - 1].idx : -1;
      internal.push_back({a, b});
    }
    s0 = s1;
  }
  // internal faces are numbered in discovery order: by their first slot
  std::sort(internal.begin(), internal.end());
  const int64_t nf = nb + (int64_t)internal.size();
  g.nface = nf;
  g.owner.assign(nf, -1);
  g.nbr.assign(nf, -1);
  g.btype.assign(nf, kInternal);
  g.S.assign(nf * nd, 0);
  g.n.assign(nf * nd, 0);
  g.area.assign(nf, 0);
  g.center.assign(nf * nd, 0);

  for (int64_t b = 0; b < nb; b++) {
    int bt = in.btype[b];
    if (bt != kSlipWall && bt != kFreestream) throw std::runtime_error("unsupported boundary type");
    g.btype[b] = bt;
    face_metrics(X, nd, &in.bfaces[b * nd], &g.S[b * nd], &g.n[b * nd], &g.area[b], &g.center[b * nd]);
    std::sort(bslots[b].begin(), bslots[b].end());   // slot order = file order of cells
    for (int64_t s : bslots[b]) {
      int64_t e = s / k;
      int m = (int)(s % k);
      int64_t fv[3];
      local_face(&in.cells[e * k], nd, m, fv);
      double out[3];
      if (nd == 2) {
        const double *a = X + fv[0] * 2, *c = X + fv[1] * 2;
        out[0] = c[1] - a[1];
        out[1] = -c[0] + a[0];
      } else {
        cross3(out, X + fv[0] * 3, X + fv[1] * 3, X + fv[2] * 3);
      }
      double dp = 0;
      for (int i = 0; i < nd; i++) dp += g.S[b * nd + i] * out[i];
      if (dp < 0)
        for (int i = 0; i < nd; i++) {
          g.S[b * nd + i] = -g.S[b * nd + i];
          g.n[b * nd + i] = -g.n[b * nd + i];
        }
      g.owner[b] = e;
    }
  }
  for (size_t q = 0; q < internal.size(); q++) {
    int64_t f = nb + (int64_t)q;
    int64_t sa = internal[q].first, sb = internal[q].second;
    int64_t e = sa / k;
    int64_t fv[3];
    local_face(&in.cells[e * k], nd, (int)(sa % k), fv);
    face_metrics(X, nd, fv, &g.S[f * nd], &g.n[f * nd], &g.area[f], &g.center[f * nd]);
    g.owner[f] = e;
    g.cface[sa] = f;
    if (sb >= 0) {
      g.nbr[f] = sb / k;
      g.cface[sb] = f;
    }
  }
  for (int64_t f = 0; f < nf; f++)
    if (g.owner[f] < 0) throw std::runtime_error("boundary face not attached to any cell");

  // ---- cell metrics ----
  g.vol.assign(ne, 0);
  g.centroid.assign(ne * nd, 0);
  g.r.assign(ns * nd, 0);
  g.cadj.assign(ns, -1);
  g.w.assign(ns, 0);
  g.rsel.assign(ns, 0);
  g.sproj.assign(ne * nd, 0);
  for (int64_t e = 0; e < ne; e++) {
    const int64_t *cv = &in.cells[e * k];
    double *cen = &g.centroid[e * nd];
    if (nd == 2) {
      const double *a = X + cv[0] * 2, *b = X + cv[1] * 2, *c = X + cv[2] * 2;
      for (int i = 0; i < 2; i++) cen[i] = (a[i] + b[i] + c[i]) / 3;
      g.vol[e] = 0.5 * std::fabs((a[0] - b[0]) * (c[1] - b[1]) - (c[0] - b[0]) * (a[1] - b[1]));
    } else {
      const double *p[4] = {X + cv[0] * 3, X + cv[1] * 3, X + cv[2] * 3, X + cv[3] * 3};
      for (int i = 0; i < 3; i++)
        cen[i] = (mode & 2) ? (p[0][i] + p[1][i] + p[2][i] + p[3][i]) / 4
                            : (p[0][i] + p[1][i] + p[2][i] + p[2][i]) / 4;
      double vol = 0;
      for (int f = 0; f < 4; f++) {
        double s[3], fc[3];
        cross3(s, p[kTetFace[f][0]], p[kTetFace[f][1]], p[kTetFace[f][2]]);
        for (int i = 0; i < 3; i++) {
          s[i] /= 2.0;
          fc[i] = 0;
          for (int q = 0; q < 3; q++) fc[i] += p[kTetFace[f][q]][i];
          fc[i] /= 3.0;
        }
        double dp = 0;
        for (int i = 0; i < 3; i++) dp += fc[i] * s[i];
        vol += dp;
      }
      g.vol[e] = (mode & 1) ? vol / 3.0 : vol;
    }
    for (int m = 0; m < k; m++) {
      int64_t f = g.cface[e * k + m];
      for (int i = 0; i < nd; i++) g.r[(e * k + m) * nd + i] = g.center[f * nd + i] - cen[i];
    }
  }
  for (int64_t e = 0; e < ne; e++) {
    for (int m = 0; m < k; m++) {
      int64_t s = e * k + m, f = g.cface[s];
      if (g.btype[f] == kInternal) g.cadj[s] = (g.owner[f] != e) ? g.owner[f] : g.nbr[f];
      int64_t j = g.cadj[s];
      if (j >= 0) {
        double dO = dist_pow(&g.centroid[e * nd], &g.center[f * nd], nd);
        double dN = dist_pow(&g.centroid[j * nd], &g.center[f * nd], nd);
        g.w[s] = dN / (dO + dN);
      }
      g.rsel[s] = (mode & 4) ? m : (f == 0 ? k - 1 : 0);
      for (int i = 0; i < nd; i++) g.sproj[e * nd + i] += std::fabs(g.S[f * nd + i]);
    }
    for (int i = 0; i < nd; i++) g.sproj[e * nd + i] /= 2.0;
  }
  return g;
}

// ---------------------------------------------------------------------------
// MSH2 reader, the subset accepted by Mesh.cpp:95-558: $PhysicalNames before
// $Nodes (the dimension is the largest physical dimension), element types
// 1/2/4, physical tag first.  Boundary names map through
// BoundaryTypeStringMap (Options.cpp:22-26).
// ---------------------------------------------------------------------------
MeshInput read_msh2(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open mesh file " + path);
  MeshInput in;
  std::string line;
  std::vector<std::pair<int, std::string>> names;
  int nd = 0;
  auto name_of = [&](int tag) -> std::string {
    for (auto &p : names) if (p.first == tag) return p.second;
    return std::string();
  };
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line == "$PhysicalNames") {
      std::getline(f, line);
      long n = std::stol(line);
      for (long i = 0; i < n; i++) {
        std::getline(f, line);
        std::istringstream ss(line);
        int d, tag;
        std::string nm;
        ss >> d >> tag >> nm;
        if (nm.size() >= 2 && nm.front() == '"') nm = nm.substr(1, nm.size() - 2);
        names.push_back({tag, nm});
        nd = std::max(nd, d);
      }
    } else if (line == "$Nodes") {
      if (nd != 2 && nd != 3) throw std::runtime_error("$PhysicalNames must precede $Nodes");
      in.ndim = nd;
      std::getline(f, line);
      long n = std::stol(line);
      in.coords.resize(n * nd);
      for (long i = 0; i < n; i++) {
        std::getline(f, line);
        const char *p = line.c_str();
        char *end;
        std::strtol(p, &end, 10);
        p = end;
        double xyz[3];
        for (int d = 0; d < 3; d++) { xyz[d] = std::strtod(p, &end); p = end; }
        for (int d = 0; d < nd; d++) in.coords[i * nd + d] = xyz[d];
      }
    } else if (line == "$Elements") {
      std::getline(f, line);
      long n = std::stol(line);
      for (long i = 0; i < n; i++) {
        std::getline(f, line);
        const char *p = line.c_str();
        char *end;
        long v[16];
        int cnt = 0;
        while (cnt < 16) {
          long x = std::strtol(p, &end, 10);
          if (end == p) break;
          v[cnt++] = x;
          p = end;
        }
        if (cnt < 3) throw std::runtime_error("malformed element line");
        int et = (int)v[1], ntag = (int)v[2];
        int base = 3 + ntag;
        if ((nd == 2 && et == 1) || (nd == 3 && et == 2)) {
          for (int q = 0; q < nd; q++) in.bfaces.push_back(v[base + q] - 1);
          std::string nm = name_of((int)v[3]);
          in.btype.push_back(nm == "SLIP_WALL" ? kSlipWall : (nm == "FREESTREAM" ? kFreestream : 0));
        } else if ((nd == 2 && et == 2) || (nd == 3 && et == 4)) {
          for (int q = 0; q < nd + 1; q++) in.cells.push_back(v[base + q] - 1);
        }
      }
    }
  }
  if (in.ndim == 0 || in.cells.empty()) throw std::runtime_error("no cells in mesh file " + path);
  return in;
}

// Morton (Z-order) of the cell centroids on a 2^21 grid per axis.
std::vector<int64_t> locality_order(const Geometry &g) {
  const int nd = g.nd;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t e = 0; e < g.ncell; e++)
    for (int i = 0; i < nd; i++) {
      lo[i] = std::min(lo[i], g.centroid[e * nd + i]);
      hi[i] = std::max(hi[i], g.centroid[e * nd + i]);
    }
  auto spread = [](uint64_t x, int nd_) {
    uint64_t r = 0;
    for (int b = 0; b < 21; b++) r |= ((x >> b) & 1ull) << (b * nd_);
    return r;
  };
  std::vector<std::pair<uint64_t, int64_t>> code(g.ncell);
  for (int64_t e = 0; e < g.ncell; e++) {
    uint64_t c = 0;
    for (int i = 0; i < nd; i++) {
      double t = (hi[i] > lo[i]) ? (g.centroid[e * nd + i] - lo[i]) / (hi[i] - lo[i]) : 0.0;
      uint64_t q = (uint64_t)(t * ((1 << 21) - 1));
      c |= spread(q, nd) << i;
    }
    code[e] = {c, e};
  }
  std::sort(code.begin(), code.end());
  std::vector<int64_t> perm(g.ncell);
  for (int64_t i = 0; i < g.ncell; i++) perm[i] = code[i].second;
  return perm;
}

}  // namespace bes
