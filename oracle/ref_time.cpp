// Timing driver for the reference's explicit path (calibration of the oracle).
//
// TEST INFRASTRUCTURE ONLY.  Linked against the reference's own Boost-free
// translation units (Config, Mesh, SolutionSnapshot, GradientGreenGauss,
// LimiterVenkatakrishnan, TimeStepCalculatorLocal; compiled from the reference
// tree by oracle/Makefile into oracle/_ref/, with their OpenMP loops) and the
// reference's flux templates.  The pseudo-time step itself lives in
// SolverExplicit.cpp, which cannot be compiled here (its header pulls in
// AMGCL, which needs Boost), so its loop is restated below with the same
// OpenMP structure:
//   * dt for all cells              SolverExplicit.cpp:269-271
//   * per stage: gradient, limiter  :279-283
//   * face loop, omp parallel for, private flux calculators, omp atomic
//     scatter to owner and neighbour :285-344
//   * element update, omp parallel for, Linf of the last stage under
//     omp critical                  :352-378
//   * face reconstruction           Solver.cpp:69-121
// It prints one JSON line: cells, iterations, threads, wall time and the last
// density Linf, and writes the final state [ncell][nv] to <state_out> (if
// given) so the caller can check it against the oracle's.
//
// Usage: bes_ref_time <config.besi> <niter> [state_out]

#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "Config.hpp"
#include "FluxFreestream.hpp"
#include "FluxRoe.hpp"
#include "FluxSlipWall.hpp"
#include "GasPerfect.hpp"
#include "GradientGreenGauss.hpp"
#include "LimiterVenkatakrishnan.hpp"
#include "Mesh.hpp"
#include "SolutionSnapshot.hpp"
#include "TimeStepCalculatorLocal.hpp"

namespace {

// one thread's flux calculators and buffers (the reference's
// PrivateVariableWrapper, copied per thread by firstprivate)
struct ThreadFlux {
  FluxRoe roe;
  FluxFreestream fs;
  FluxSlipWall sw;
  double L[VAR_BUFFER], R[VAR_BUFFER], F[VAR_BUFFER];
};

// U + (phi*grad).r of one side (Solver.cpp:69-121)
void side_state(Mesh *mesh, SolutionSnapshot *s, Gradient *g, Limiter *lim, Element *e,
                unsigned long iFace, int nv, int nd, double *out) {
  const unsigned long id = e->getId();
  for (int v = 0; v < nv; v++) {
    const double phi = lim->getLimiterValue(id, v);
    out[v] = s->getConservedVariable(id, v);
    for (int d = 0; d < nd; d++)
      out[v] += phi * g->getGradient(id, v, d) * e->getVectorToFaceFromFaceId(iFace, d);
  }
  (void)mesh;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s <config.besi> <niter> [state_out]\n", argv[0]);
    return 2;
  }
  const long niter = atol(argv[2]);
  Config cfg;
  Mesh mesh;
  GasPerfect gas;
  cfg.initialize(argv[1]);
  mesh.initialize(&cfg);
  cfg.initializePostMesh(&mesh);
  gas.initialize(&cfg);
  const int nd = mesh.getNumDim(), nv = cfg.getNumConservedVariable();
  const unsigned long nf = mesh.getNumFace(), ne = mesh.getNumElement();
  double *finf = cfg.getFreestreamPrimitiveVariables(0);
  std::vector<double> u0(nv);
  gas.computeAllConservedVariable(finf, u0.data());

  SolutionSnapshot s0, s1;
  s0.initialize(&cfg, &mesh, &gas);
  s1.initialize(&cfg, &mesh, &gas);
  for (unsigned long e = 0; e < ne; e++) {
    s0.setConservedVariable(e, u0.data());
    s1.setConservedVariable(e, u0.data());
  }
  GradientGreenGauss grad;
  grad.initialize(&cfg, &mesh, VariableType::CONSERVED);
  LimiterVenkatakrishnan lim;
  lim.initialize(&cfg, &mesh, VariableType::CONSERVED);
  TimeStepCalculatorLocal tsl;
  tsl.initialize(&cfg, &gas, &mesh);

  const int nthreads = omp_get_max_threads();
  std::vector<ThreadFlux> tf(nthreads);
  for (auto &t : tf) {
    t.roe.initialize(&cfg, &gas);
    t.fs.initialize(&cfg, &gas);
    t.sw.initialize(&cfg, &gas);
  }
  std::vector<double> sub(ne * nv);
  std::vector<double> linf(nv);
  SolutionSnapshot *cur0 = &s0, *cur1 = &s1;
  const int nsub = cfg.getNumSubTimeStep();

  const auto t0 = std::chrono::steady_clock::now();
  for (long it = 0; it < niter; it++) {
    tsl.computeAllTimeStep(cur0, cfg.getCourantNum());
    for (int iSub = 0; iSub < nsub; iSub++) {
      std::fill(sub.begin(), sub.end(), 0.0);
      SolutionSnapshot *ss = iSub == 0 ? cur0 : cur1;
      grad.computeGradient(ss);
      lim.computeLimiterTerm(ss, &grad);
#pragma omp parallel for default(shared)
      for (unsigned long f = 0; f < nf; f++) {
        ThreadFlux &t = tf[omp_get_thread_num()];
        Face *fc = mesh.getFace(f);
        side_state(&mesh, ss, &grad, &lim, fc->getOwnerElement(), f, nv, nd, t.L);
        Element *nb = fc->getNeighborElement();
        if (nb) side_state(&mesh, ss, &grad, &lim, nb, f, nv, nd, t.R);
        const BoundaryType bt = fc->getBoundaryType();
        if (bt == BoundaryType::INTERNAL) t.roe.computeFlux(t.L, t.R, fc->getNonDimNormal(), t.F);
        else if (bt == BoundaryType::FREESTREAM) t.fs.computeFlux(t.L, finf, fc->getNonDimNormal(), t.F);
        else t.sw.computeFlux(t.L, NULL, fc->getNonDimNormal(), t.F);
        const unsigned long o = fc->getOwnerElement()->getId();
        for (int v = 0; v < nv; v++) {
          if (t.F[v] != t.F[v]) {
#pragma omp critical
            {
              fprintf(stderr, "NaN flux at face %lu\n", f);
              exit(1);
            }
          }
#pragma omp atomic
          sub[o * nv + v] += t.F[v] * fc->getArea();
        }
        if (nb) {
          const unsigned long j = nb->getId();
          for (int v = 0; v < nv; v++) {
#pragma omp atomic
            sub[j * nv + v] -= t.F[v] * fc->getArea();
          }
        }
      }
      const bool last = iSub == nsub - 1;
      if (last) std::fill(linf.begin(), linf.end(), 0.0);
#pragma omp parallel for default(shared)
      for (unsigned long e = 0; e < ne; e++) {
        for (int v = 0; v < nv; v++) {
          if (last) {
#pragma omp critical
            linf[v] = std::abs(sub[e * nv + v]) > linf[v] ? std::abs(sub[e * nv + v]) : linf[v];
          }
          const double nu = cur0->getConservedVariable(e, v) -
                            (cfg.getTimeStepConstant(iSub) * tsl.getTimeStep(e) / mesh.getElement(e)->getVolume()) *
                                sub[e * nv + v];
          cur1->setConservedVariable(e, v, nu);
        }
      }
    }
    std::swap(cur0, cur1);
  }
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (argc > 3) {
    FILE *f = fopen(argv[3], "wb");
    if (!f) { perror(argv[3]); return 2; }
    for (unsigned long e = 0; e < ne; e++)
      for (int v = 0; v < nv; v++) {
        const double x = cur0->getConservedVariable(e, v);
        fwrite(&x, sizeof(double), 1, f);
      }
    fclose(f);
  }
  printf("{\"cells\": %lu, \"iterations\": %ld, \"threads\": %d, \"wall_s\": %.6f, "
         "\"cell_updates_per_s\": %.6e, \"density_linf_last\": %.17g}\n",
         ne, niter, nthreads, wall, (double)ne * niter / wall, linf[0]);
  return 0;
}
