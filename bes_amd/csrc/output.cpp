// Solution files in the reference's formats: legacy VTK and CSV of the cell
// primitives, and the slip-wall CSV (SolutionSnapshot.cpp:146-308), with the
// reference's dimensional conversion (GasPerfect.tpp:34-111,294-320).  Host
// code: the state comes down once per file.
#include <hip/hip_runtime.h>

#include <cmath>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/bes_hip.h"
#include "context.hpp"

namespace bes {

int state_to_host(bes_ctx *c, const double *dsoa, int ncomp, double *host);

// primitive [rho, u_d, p, T] of one conserved state, then the reference's
// "dimensional" form [rho, u_d / c, gamma p, gamma R T]
static void dimensional_primitive(const GasConst &g, int nd, const double *U, double *P) {
  const double rho = U[0], rhoE = U[nd + 1];
  double q2 = 0;
  for (int d = 0; d < nd; d++) q2 += U[1 + d] * U[1 + d];
  const double p = (g.gamma - 1.0) * (rhoE - q2 / (2.0 * rho));
  const double T = (1.0 / (g.cv * rho)) * (rhoE - q2 / (2.0 * rho));
  const double c = std::sqrt(g.gamma * g.R * T);
  P[0] = rho;
  for (int d = 0; d < nd; d++) P[1 + d] = (U[1 + d] / rho) / c;
  P[nd + 1] = g.gamma * p;
  P[nd + 2] = g.gamma * g.R * T;
}

static int host_state(bes_ctx *c, int cond, std::vector<double> &U) {
  const size_t N = c->N, nv = c->nv;
  std::vector<double> all(N * nv * c->B);
  int s = state_to_host(c, c->U0, c->nv, all.data());
  if (s) return s;
  U.assign(all.begin() + (size_t)cond * N * nv, all.begin() + (size_t)(cond + 1) * N * nv);
  return BES_OK;
}

}  // namespace bes

using namespace bes;

extern "C" {

int bes_write_solution(bes_ctx *c, const char *path, int format, int cond) {
  if (!c || !path || (format != 0 && format != 1) || cond < 0 || cond >= c->B) {
    set_error("invalid argument (format 0 VTK, 1 CSV; condition index)");
    return BES_ERR_ARG;
  }
  std::vector<double> U;
  int s = host_state(c, cond, U);
  if (s) return s;
  const int nd = c->nd, nv = c->nv, np = nd + 3;
  const int64_t N = c->N;
  std::vector<double> P((size_t)N * np);
  for (int64_t e = 0; e < N; e++) dimensional_primitive(c->gc, nd, &U[e * nv], &P[e * np]);
  std::ofstream out(path);
  if (!out) { set_error(std::string("cannot open ") + path); return BES_ERR_ARG; }
  if (format == 1) {   // SolutionSnapshot.cpp:146-170: centroid, then the primitives
    for (int64_t e = 0; e < N; e++) {
      for (int d = 0; d < nd; d++) out << c->geo.centroid[e * nd + d] << ",";
      for (int q = 0; q < np; q++) out << P[e * np + q] << ",";
      out << std::endl;
    }
    return BES_OK;
  }
  // SolutionSnapshot.cpp:172-266
  const int64_t nnode = (int64_t)c->node_xyz.size() / nd;
  const int kv = nd + 1;
  out << "# vtk DataFile Version 2.0" << std::endl << path << std::endl << "ASCII" << std::endl
      << "DATASET UNSTRUCTURED_GRID" << std::endl << "POINTS " << nnode << " float" << std::endl;
  for (int64_t v = 0; v < nnode; v++) {
    for (int d = 0; d < nd; d++) out << c->node_xyz[v * nd + d] << " ";
    if (nd == 2) out << 0.0 << " ";
    out << std::endl;
  }
  out << "CELLS " << N << " " << N * (kv + 1) << std::endl;
  for (int64_t e = 0; e < N; e++) {
    out << kv << " ";
    for (int q = 0; q < kv; q++) out << c->cell_nodes[e * kv + q] << " ";
    out << std::endl;
  }
  out << "CELL_TYPES " << N << std::endl;
  for (int64_t e = 0; e < N; e++) out << (nd == 2 ? 5 : 10) << std::endl;
  out << "CELL_DATA " << N << std::endl;
  const char *names[3] = {"rho_nondim", "P_nondim", "T_nondim"};
  const int cols[3] = {0, nd + 1, nd + 2};
  for (int a = 0; a < 3; a++) {
    out << "SCALARS " << names[a] << " double" << std::endl << "LOOKUP_TABLE default" << std::endl;
    for (int64_t e = 0; e < N; e++) out << P[e * np + cols[a]] << std::endl;
  }
  out << "VECTORS M double" << std::endl;
  for (int64_t e = 0; e < N; e++) {
    for (int d = 0; d < nd; d++) out << P[e * np + 1 + d] << " ";
    if (nd == 2) out << 0.0 << " ";
    out << std::endl;
  }
  return BES_OK;
}

// SolutionSnapshot.cpp:268-308: per slip-wall face its centre, area vector and
// the owner cell's primitives (zeroth-order surface values)
int bes_write_slip_wall(bes_ctx *c, const char *path, int cond) {
  if (!c || !path || cond < 0 || cond >= c->B) { set_error("invalid argument"); return BES_ERR_ARG; }
  std::vector<double> U;
  int s = host_state(c, cond, U);
  if (s) return s;
  const int nd = c->nd, nv = c->nv, np = nd + 3;
  std::ofstream out(path);
  if (!out) { set_error(std::string("cannot open ") + path); return BES_ERR_ARG; }
  std::vector<double> P(np);
  const Geometry &g = c->geo;
  for (int64_t f = 0; f < g.nface; f++) {
    if (g.btype[f] != kSlipWall) continue;
    const int64_t e = g.owner[f];
    dimensional_primitive(c->gc, nd, &U[e * nv], P.data());
    for (int d = 0; d < nd; d++) out << g.center[f * nd + d] << ",";
    for (int d = 0; d < nd; d++) out << g.S[f * nd + d] << ",";
    for (int q = 0; q < np; q++) out << P[q] << ",";
    out << std::endl;
  }
  return BES_OK;
}

}  // extern "C"
