// bes_hip: command-line driver with the reference's usage (BES.cpp:40-121):
//   bes_hip <config.besi>
// reads the configuration and the MSH2 mesh it names, runs the explicit or
// implicit solver on the GPU and writes OUTPUT_FOLDER/residual.csv in the
// reference's format.  Built on the C++ classes of include/bes/bes.hpp.
#include <cstdio>
#include <memory>

#include "../../include/bes/bes.hpp"

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <config.besi>\n", argv[0]);
    return 2;
  }
  try {
    bes_hip::Config config;
    bes_hip::Mesh mesh;
    bes_hip::GasPerfect gas;
    config.initialize(argv[1]);
    mesh.initialize(&config);
    config.initializePostMesh(mesh.getNumDim());
    gas.initialize(&config);
    std::unique_ptr<bes_hip::Solver> solver;
    if (config.getSolverType() == bes_hip::SolverType::IMPLICIT) solver.reset(new bes_hip::SolverImplicit());
    else solver.reset(new bes_hip::SolverExplicit());
    solver->initialize(&config, &gas, &mesh);
    solver->setUniversalInitialCondition(config.getFreestreamPrimitiveVariables(0));
    solver->solve();
    std::printf("%lu cells, %zu pseudo-time steps\n", mesh.getNumElement(), solver->history().size());
  } catch (const bes_hip::Error &e) {
    std::fprintf(stderr, "bes_hip: %s (status %d, face %ld)\n", e.what(), e.status, e.face);
    return 1;
  }
  return 0;
}
