// bes.hpp — C++ host classes of the MI355X BES solver core.
//
// The reference (joel-ho/BES) exposes its residual and implicit-update path as
// C++ classes; these classes keep the reference's names, method signatures,
// argument meanings and conventions, and run on the GPU through the C ABI of
// bes_hip.h.  Pass-level methods (computeGradient, computeLimiterTerm,
// computeAllTimeStep, solve) launch device work; per-element getters
// (getGradient, getLimiterValue, getTimeStep, getConservedVariable) read a
// host mirror that is refreshed lazily after each pass.  Errors throw
// bes_hip::Error instead of calling exit() as the reference does.
//
// Interfaces mirrored (reference file:line):
//   Config           Config.hpp:35-148, parser Config.cpp:116-422
//   Mesh             Mesh.hpp:39-84 (Mesh::initialize Mesh.cpp:66-93)
//   GasPerfect       GasPerfect.hpp:27-78 (initialize only: the gas model runs on the device)
//   SolutionSnapshot SolutionSnapshot.hpp:34-83 (conserved-variable accessors)
//   Gradient, GradientGreenGauss        Gradient.hpp:28-51
//   Limiter, LimiterVenkatakrishnan     Limiter.hpp:28-53
//   TimeStepCalculator{,Local,Global}   TimeStepCalculator.hpp:28-48
//   Flux_, FluxRoe, FluxFreestream, FluxSlipWall   Flux.hpp:26-51
//   FluxJacobian and its Roe/Freestream/SlipWall forms   FluxJacobian.hpp:26-46
//   CoefficientMatrix                   CoefficientMatrix.hpp:51-87
//   Solver, SolverExplicit, SolverImplicit   Solver.hpp:36-74, SolverExplicit.hpp:60-81,
//                                            SolverImplicit.hpp:81-119
#pragma once

#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../bes_hip.h"

namespace bes_hip {

class Error : public std::runtime_error {
 public:
  int status;
  long face;
  Error(int s, const std::string &msg, long f = -1) : std::runtime_error(msg), status(s), face(f) {}
};

inline void check(int s, const bes_ctx *c = nullptr) {
  if (s != BES_OK) {
    std::string m = bes_last_error_message();
    throw Error(s, m.empty() ? bes_status_string(s) : m, c ? bes_last_error_face(c) : -1);
  }
}

enum class SolverType { EXPLICIT = 1, IMPLICIT };
enum class LocalTimeStep { YES = 1, NO };
enum class AdaptiveCourantNum { YES = 1, NO };
enum class LinearSolver { FGMRES = 1, GMRES, LGMRES, BICGSTAB };   // Options.hpp:67
enum VariableType { CONSERVED = 0, PRIMITIVE, COUNT };

// ---------------------------------------------------------------------------
// Config: the .besi keys of the solver path (Config.cpp:73-114 defaults,
// 116-422 parser: "KEY = v1, v2 // comment", unknown keys are errors).
// ---------------------------------------------------------------------------
class Mesh;

class Config {
 public:
  void initialize(const std::string &path) {
    setDefaultValues_();
    std::ifstream f(path);
    if (!f) throw Error(BES_ERR_ARG, "cannot open config file " + path);
    std::string line;
    while (std::getline(f, line)) {
      auto c = line.find("//");
      if (c != std::string::npos) line = line.substr(0, c);
      std::string s;
      for (char ch : line) if (ch != ' ' && ch != '\t' && ch != '\r') s += ch;
      if (s.empty()) continue;
      auto eq = s.find('=');
      if (eq == std::string::npos) throw Error(BES_ERR_ARG, "malformed config line: " + line);
      std::string key = s.substr(0, eq);
      std::vector<std::string> arg;
      std::stringstream ss(s.substr(eq + 1));
      std::string tok;
      while (std::getline(ss, tok, ',')) arg.push_back(tok);
      if (arg.empty()) arg.push_back("");
      apply_(key, arg);
    }
    nCurrentSolution_ = solverType_ == SolverType::EXPLICIT ? 2 : 1;   // Config.cpp:430-446
  }
  // Config.cpp:455-515: variable counts and the non-dimensional freestream
  void initializePostMesh(int nDim) {
    nDim_ = nDim;
    nConserved_ = nDim + 2;
    nPrimitive_ = nDim + 3;
    freestream_.assign(nPrimitive_, 0.0);
    freestream_[0] = 1.0;
    for (int d = 0; d < nDim; d++) freestream_[d + 1] = mach_.size() > (size_t)d ? mach_[d] : 0.0;
    freestream_[nPrimitive_ - 2] = 1.0 / gamma_;
    freestream_[nPrimitive_ - 1] = 1.0 / (gamma_ * R_);
    ramp_ = freestream_;
  }
  // Config.cpp:688-706 (the ramp scales the velocity only)
  double *getFreestreamPrimitiveVariables(unsigned long iTimeStep) {
    if (iTimeStep >= nRamp_) return freestream_.data();
    for (int d = 0; d < nDim_; d++)
      ramp_[d + 1] = freestream_[d + 1] * (rampStart_ + double(iTimeStep) / double(nRamp_) * (1 - rampStart_));
    return ramp_.data();
  }
  const std::string &getMeshFilePath() const { return meshFile_; }
  const std::string &getOutputFolderPath() const { return outputFolder_; }
  int getOutputFormat() const { return outputFormat_; }          // 0 VTK, 1 CSV
  bool getOutputSlipWall() const { return outputSlipWall_; }
  short getNumDim() const { return (short)nDim_; }
  short getNumConservedVariable() const { return (short)nConserved_; }
  short getNumPrimitiveVariable() const { return (short)nPrimitive_; }
  SolverType getSolverType() const { return solverType_; }
  LocalTimeStep getLocalTimeStepOption() const { return localTimeStep_; }
  unsigned long getNumTimeStep() const { return nTimeStep_; }
  unsigned long getNumTimeStepBetweenSave() const { return nSave_; }
  short getNumSubTimeStep() const { return (short)alpha_.size(); }
  double getTimeStepConstant(short i) const { return alpha_[i]; }
  const std::vector<double> &getTimeStepConstants() const { return alpha_; }
  double getCourantNum() const { return cfl_; }
  double getAdaptiveCourantNumMax() const { return cflMax_; }
  const double *getAdaptiveCourantNumExponent() const { return exponent_; }
  AdaptiveCourantNum getAdaptiveCourantNumOption() const {   // Config.cpp:631-634
    return cflMax_ > cfl_ ? AdaptiveCourantNum::YES : AdaptiveCourantNum::NO;
  }
  short getNumCurrentSolution() const { return (short)nCurrentSolution_; }
  double getVenkatakrishnanLimiterConstant() const { return venkat_; }
  double getHartenCorrectionMachNum() const { return harten_; }
  double getGasConstant() const { return R_; }
  double getRatioSpecificHeat() const { return gamma_; }
  LinearSolver getLinearSolverType() const { return linearSolver_; }
  int getLinearSolverNumIter() const { return linMaxIter_; }
  double getLinearSolverTol() const { return linTol_; }
  bes_gas gas() const { return bes_gas{gamma_, R_, harten_, venkat_}; }
  // build-specific keys (no counterpart in the reference)
  int getGeometryMode() const { return geomMode_; }
  int getPrecision() const { return precision_; }
  int getPreconditioner() const { return precond_; }
  bool getAmgAggregation() const { return amgAggregation_; }
  int getAmgRelax() const { return amgRelax_; }   // 0 damped Jacobi, 1 SPAI0, 2 block Jacobi, -1 other
  int getBlockRepeat() const { return blockRepeat_; }
  double getLimiterFreezeDrop() const { return limiterFreeze_; }

 private:
  void setDefaultValues_() {   // Config.cpp:73-114
    solverType_ = SolverType::EXPLICIT;
    localTimeStep_ = LocalTimeStep::NO;
    nRamp_ = 0;
    rampStart_ = 0.5;
    cfl_ = 2.0;
    cflMax_ = 0;
    exponent_[0] = exponent_[1] = 1;
    alpha_ = {0.0833, 0.2069, 0.4265, 1.0};
    linearSolver_ = LinearSolver::FGMRES;
    linMaxIter_ = 1000;
    linTol_ = 1e-7;
    venkat_ = 5.0;
    harten_ = 0.05;
    R_ = 287;
    gamma_ = 1.4;
  }
  static double num(const std::string &s) { return std::stod(s); }
  void apply_(const std::string &k, const std::vector<std::string> &a) {
    static const char *ignored[] = {"MESH_FORMAT", "EQUATION",
                                    "EXPLICIT_SCHEME", "GRADIENT_SCHEME", "LIMITER_SCHEME", "FLUX_SCHEME"};
    for (const char *ig : ignored) if (k == ig) return;   // single scheme each on this path
    if (k == "MESH_FILE") meshFile_ = a[0];
    else if (k == "OUTPUT_FOLDER") outputFolder_ = a[0];
    else if (k == "OUTPUT_FORMAT") outputFormat_ = a[0] == "CSV" ? 1 : 0;
    else if (k == "OUTPUT_SLIP_WALL") outputSlipWall_ = a[0] == "YES";
    else if (k == "SOLVER_TYPE") solverType_ = a[0] == "IMPLICIT" ? SolverType::IMPLICIT : SolverType::EXPLICIT;
    else if (k == "MULTI_TIME_STEP_OPTIONS") {
      int n = (int)num(a[0]);
      alpha_.assign(n, 0.0);
      for (int i = 0; i < n && i + 1 < (int)a.size(); i++) alpha_[i] = num(a[i + 1]);
    } else if (k == "COURANT_NUM") cfl_ = num(a[0]);
    else if (k == "ADAPTIVE_COURANT_NUM_MAX") cflMax_ = num(a[0]);
    else if (k == "ADAPTIVE_COURANT_NUM_EXPONENT") {
      exponent_[0] = num(a[0]);
      if (a.size() > 1) exponent_[1] = num(a[1]);
    } else if (k == "LOCAL_TIME_STEP") localTimeStep_ = a[0] == "YES" ? LocalTimeStep::YES : LocalTimeStep::NO;
    else if (k == "NUMBER_OF_TIME_STEP") nTimeStep_ = (unsigned long)num(a[0]);
    else if (k == "SAVE_INTERVAL") nSave_ = (unsigned long)num(a[0]);
    else if (k == "FREESTREAM_RAMP") nRamp_ = (unsigned long)num(a[0]);
    else if (k == "FREESTREAM_RAMP_START_FRAC") rampStart_ = num(a[0]);
    else if (k == "VENKATAKRISHNAN_FACTOR") venkat_ = num(a[0]);
    else if (k == "HARTEN_CORRECTION_LIMIT") harten_ = num(a[0]);
    else if (k == "GAS_CONSTANT") R_ = num(a[0]);
    else if (k == "SPECIFIC_HEAT_RATIO") gamma_ = num(a[0]);
    else if (k == "FREESTREAM_MACH") { mach_.clear(); for (auto &x : a) mach_.push_back(num(x)); }
    else if (k == "LINEAR_SOLVER") {   // Config.cpp:345-358; other values keep the default
      if (a[0] == "FGMRES") linearSolver_ = LinearSolver::FGMRES;
      else if (a[0] == "GMRES") linearSolver_ = LinearSolver::GMRES;
      else if (a[0] == "LGMRES") linearSolver_ = LinearSolver::LGMRES;
      else if (a[0] == "BICGSTAB") linearSolver_ = LinearSolver::BICGSTAB;
    } else if (k == "LINEAR_SOLVER_MAX_ITER") linMaxIter_ = (int)num(a[0]);
    else if (k == "LINEAR_SOLVER_ABS_TOL") linTol_ = num(a[0]);
    // keys of this build
    else if (k == "GEOMETRY_MODE") geomMode_ = a[0] == "EXACT" ? BES_GEOM_EXACT : BES_GEOM_COMPAT;
    else if (k == "LINEAR_SOLVER_PRECISION") precision_ = (int)num(a[0]);
    else if (k == "PRECONDITIONER")
      precond_ = a[0] == "NONE" ? 0 : a[0] == "BLOCK_JACOBI" ? 1 : a[0] == "LEARNED" ? 3 : a[0] == "AMG" ? 4 : 2;
    // Config.cpp:365-400; used when PRECONDITIONER = AMG
    else if (k == "AMG_PRECOND_COARSENER") amgAggregation_ = a[0] == "AGGREGATION";
    else if (k == "AMG_PRECOND_SMOOTHER")
      amgRelax_ = a[0] == "DAMPED_JACOBI" ? 0 : a[0] == "SPAI0" ? 1 : a[0] == "BLOCK_JACOBI" ? 2 : -1;
    else if (k == "JACOBIAN_BLOCK_REPEAT") blockRepeat_ = a[0] == "NO" ? 0 : 1;
    else if (k == "LIMITER_FREEZE_DROP") limiterFreeze_ = num(a[0]);   // 0: never (bes_set_limiter_freeze)
    else throw Error(BES_ERR_ARG, "unknown config key " + k);
  }
  std::string meshFile_, outputFolder_ = ".";
  int outputFormat_ = 0;
  bool outputSlipWall_ = false;
  SolverType solverType_;
  LocalTimeStep localTimeStep_;
  unsigned long nRamp_ = 0, nTimeStep_ = 0, nSave_ = 1;
  double rampStart_ = 0.5, cfl_ = 2.0, cflMax_ = 0, exponent_[2] = {1, 1};
  std::vector<double> alpha_, mach_, freestream_, ramp_;
  LinearSolver linearSolver_ = LinearSolver::FGMRES;
  int linMaxIter_ = 1000;
  double linTol_ = 1e-7, venkat_ = 5.0, harten_ = 0.05, R_ = 287, gamma_ = 1.4;
  int nDim_ = 0, nConserved_ = 0, nPrimitive_ = 0, nCurrentSolution_ = 1;
  int geomMode_ = BES_GEOM_COMPAT, precision_ = 64, precond_ = 2, blockRepeat_ = 1;
  double limiterFreeze_ = 0.0;
  bool amgAggregation_ = true;
  int amgRelax_ = 0;
};

// ---------------------------------------------------------------------------
// Mesh: reads the MSH2 file into a device context (Mesh.cpp:66-93)
// ---------------------------------------------------------------------------
class Mesh {
 public:
  ~Mesh() { if (ctx_) bes_ctx_destroy(ctx_); }
  void initialize(Config *config) {
    bes_gas g = config->gas();
    check(bes_ctx_create_msh2(config->getMeshFilePath().c_str(), config->getGeometryMode(), &g, &ctx_));
    int nd;
    long nc, nf;
    check(bes_ctx_info(ctx_, &nd, &nc, &nf));
    nDim_ = (short)nd;
    nElement_ = (unsigned long)nc;
    nFace_ = (unsigned long)nf;
  }
  short getNumDim() const { return nDim_; }
  unsigned long getNumElement() const { return nElement_; }
  unsigned long getNumFace() const { return nFace_; }
  bes_ctx *ctx() const { return ctx_; }

 private:
  bes_ctx *ctx_ = nullptr;
  short nDim_ = 0;
  unsigned long nElement_ = 0, nFace_ = 0;
};

class GasPerfect {
 public:
  void initialize(Config *config) { config_ = config; }
  Config *config() const { return config_; }

 private:
  Config *config_ = nullptr;
};
typedef GasPerfect Gas;

// ---------------------------------------------------------------------------
// SolutionSnapshot: the device state with a host mirror
// ---------------------------------------------------------------------------
class SolutionSnapshot {
 public:
  void initialize(Config *config, Mesh *mesh, Gas *) {
    mesh_ = mesh;
    nv_ = config->getNumConservedVariable();
    host_.assign(mesh->getNumElement() * nv_, 0.0);
  }
  double getConservedVariable(unsigned long e, short v) { pull_(); return host_[e * nv_ + v]; }
  double *getConservedVariable(unsigned long e) { pull_(); return &host_[e * nv_]; }
  void setConservedVariable(unsigned long e, short v, double x) { pull_(); host_[e * nv_ + v] = x; dirty_ = true; }
  void setConservedVariable(unsigned long e, const double *x) {
    pull_();
    for (int v = 0; v < nv_; v++) host_[e * nv_ + v] = x[v];
    dirty_ = true;
  }
  void addToConservedVariable(unsigned long e, short v, double x) { pull_(); host_[e * nv_ + v] += x; dirty_ = true; }
  // device synchronisation used by the pass-level classes
  void push() { if (dirty_) { check(bes_set_state(mesh_->ctx(), host_.data())); dirty_ = false; } stale_ = false; }
  void markDeviceNewer() { stale_ = true; }
  bes_ctx *ctx() const { return mesh_->ctx(); }

 private:
  void pull_() {
    if (stale_) { check(bes_get_state(mesh_->ctx(), host_.data())); stale_ = false; }
  }
  Mesh *mesh_ = nullptr;
  int nv_ = 0;
  std::vector<double> host_;
  bool dirty_ = true, stale_ = false;
};

// ---------------------------------------------------------------------------
// Gradient / Limiter (one fused device pass; the limiter reuses it)
// ---------------------------------------------------------------------------
class Gradient {
 public:
  virtual ~Gradient() {}
  void initialize(Config *config, Mesh *mesh, VariableType) {
    nv_ = config->getNumConservedVariable();
    nd_ = mesh->getNumDim();
    host_.assign(mesh->getNumElement() * nv_ * nd_, 0.0);
  }
  virtual void computeGradient(SolutionSnapshot *s) = 0;
  double getGradient(unsigned long e, short v, short d) { pull_(); return host_[(e * nv_ + v) * nd_ + d]; }

 protected:
  void pull_() { if (stale_ && ctx_) { check(bes_get_gradient(ctx_, host_.data())); stale_ = false; } }
  bes_ctx *ctx_ = nullptr;
  int nv_ = 0, nd_ = 0;
  std::vector<double> host_;
  bool stale_ = false;
  friend class LimiterVenkatakrishnan;
};

class GradientGreenGauss : public Gradient {
 public:
  // GradientGreenGauss.cpp:21-135 (with the limiter of the same state, fused)
  void computeGradient(SolutionSnapshot *s) override {
    s->push();
    ctx_ = s->ctx();
    check(bes_compute_gradient_limiter(ctx_));
    stale_ = true;
  }
};

class Limiter {
 public:
  virtual ~Limiter() {}
  void initialize(Config *config, Mesh *mesh, VariableType) {
    nv_ = config->getNumConservedVariable();
    host_.assign(mesh->getNumElement() * nv_, 0.0);
  }
  virtual void computeLimiterTerm(SolutionSnapshot *s, Gradient *g) = 0;
  double getLimiterValue(unsigned long e, short v) {
    if (stale_ && ctx_) { check(bes_get_limiter(ctx_, host_.data())); stale_ = false; }
    return host_[e * nv_ + v];
  }

 protected:
  bes_ctx *ctx_ = nullptr;
  int nv_ = 0;
  std::vector<double> host_;
  bool stale_ = false;
};

class LimiterVenkatakrishnan : public Limiter {
 public:
  // LimiterVenkatakrishnan.cpp:21-99: computed in the gradient's device pass
  void computeLimiterTerm(SolutionSnapshot *s, Gradient *g) override {
    if (g->ctx_ != s->ctx()) g->computeGradient(s);
    ctx_ = s->ctx();
    stale_ = true;
  }
};

// ---------------------------------------------------------------------------
// Time step
// ---------------------------------------------------------------------------
class TimeStepCalculator {
 public:
  virtual ~TimeStepCalculator() {}
  virtual void initialize(Config *config, Gas *, Mesh *mesh) {
    host_.assign(mesh->getNumElement(), 0.0);
    (void)config;
  }
  virtual void computeAllTimeStep(SolutionSnapshot *s, const double &courantNum) = 0;
  virtual double getTimeStep(const unsigned long &e) {
    if (stale_) { check(bes_get_time_step(ctx_, host_.data())); stale_ = false; }
    return host_[e];
  }

 protected:
  void run_(SolutionSnapshot *s, double cfl, int local) {
    s->push();
    ctx_ = s->ctx();
    check(bes_compute_time_step(ctx_, cfl, local));
    stale_ = true;
  }
  bes_ctx *ctx_ = nullptr;
  std::vector<double> host_;
  bool stale_ = false;
};

class TimeStepCalculatorLocal : public TimeStepCalculator {
 public:
  void computeAllTimeStep(SolutionSnapshot *s, const double &cfl) override { run_(s, cfl, 1); }
};
class TimeStepCalculatorGlobal : public TimeStepCalculator {
 public:
  void computeAllTimeStep(SolutionSnapshot *s, const double &cfl) override { run_(s, cfl, 0); }
};

// ---------------------------------------------------------------------------
// Per-face fluxes and Jacobians (device batches; the per-face call is a batch of 1)
// ---------------------------------------------------------------------------
template <typename T>
class Flux_ {
 public:
  virtual ~Flux_() {}
  virtual void initialize(Config *config, Gas *) { config_ = config; }
  // left/right as the reference: conserved (Roe), freestream primitive (far
  // field) or NULL (slip wall); flux per unit area, overwritten
  virtual void computeFlux(T const *left, T const *right, double const *normal, T *flux) {
    computeFluxBatch(1, left, right, normal, flux);
  }
  void computeFluxBatch(long n, T const *left, T const *right, double const *normal, T *flux) {
    const int nv = config_->getNumConservedVariable(), nd = config_->getNumDim();
    std::vector<double> ur((size_t)n * (nv + 1), 0.0);
    const int nr = kind_ == BES_INTERNAL ? nv : (kind_ == BES_FREESTREAM ? nv + 1 : 0);
    for (long i = 0; i < n && right; i++)
      for (int q = 0; q < nr; q++) ur[i * (nv + 1) + q] = right[i * nr + q];
    bes_gas g = config_->gas();
    check(bes_flux_batch(nd, kind_, n, left, ur.data(), normal, &g, flux));
  }

 protected:
  explicit Flux_(int kind) : kind_(kind) {}
  Config *config_ = nullptr;
  int kind_;
};
class FluxRoe : public Flux_<double> { public: FluxRoe() : Flux_<double>(BES_INTERNAL) {} };
class FluxFreestream : public Flux_<double> { public: FluxFreestream() : Flux_<double>(BES_FREESTREAM) {} };
class FluxSlipWall : public Flux_<double> { public: FluxSlipWall() : Flux_<double>(BES_SLIP_WALL) {} };
typedef Flux_<double> Flux;

class FluxJacobian {
 public:
  virtual ~FluxJacobian() {}
  virtual void initialize(Config *config) { config_ = config; }
  // row-major nv x nv blocks; rightJacobian may be NULL (FluxJacobian.hpp:36-44)
  virtual void computeJacobian(double const *left, double const *right, double const *normal,
                               double *leftJacobian, double *rightJacobian) {
    const int nv = config_->getNumConservedVariable(), nd = config_->getNumDim();
    std::vector<double> ur(nv + 1, 0.0), jr(nv * nv);
    const int nr = kind_ == BES_INTERNAL ? nv : (kind_ == BES_FREESTREAM ? nv + 1 : 0);
    for (int q = 0; q < nr && right; q++) ur[q] = right[q];
    bes_gas g = config_->gas();
    check(bes_jacobian_batch(nd, kind_, 1, left, ur.data(), normal, &g, leftJacobian, jr.data()));
    if (rightJacobian) for (int q = 0; q < nv * nv; q++) rightJacobian[q] = jr[q];
  }

 protected:
  explicit FluxJacobian(int kind) : kind_(kind) {}
  Config *config_ = nullptr;
  int kind_;
};
class FluxJacobianRoe : public FluxJacobian { public: FluxJacobianRoe() : FluxJacobian(BES_INTERNAL) {} };
class FluxJacobianFreestream : public FluxJacobian { public: FluxJacobianFreestream() : FluxJacobian(BES_FREESTREAM) {} };
class FluxJacobianSlipWall : public FluxJacobian { public: FluxJacobianSlipWall() : FluxJacobian(BES_SLIP_WALL) {} };
typedef FluxJacobianRoe FluxJacobianRoeTwoDim;
typedef FluxJacobianRoe FluxJacobianRoeThreeDim;
typedef FluxJacobianFreestream FluxJacobianFreestreamTwoDim;
typedef FluxJacobianFreestream FluxJacobianFreestreamThreeDim;
typedef FluxJacobianSlipWall FluxJacobianSlipWallTwoDim;
typedef FluxJacobianSlipWall FluxJacobianSlipWallThreeDim;

// preconditioner state after bes_implicit_setup: the learned operator's
// initial weights, or the AMG parameters of the reference's test
// configurations (eps_strong 0.08, over_interp 1.5, damping 0.72, coarse_enough
// 500: CoefficientMatrix.cpp:231-280 with AMGCL's defaults)
inline void setupPreconditioner_(bes_ctx *ctx, const Config *config) {
  // Krylov method (CoefficientMatrix.cpp:282-296): FGMRES and GMRES both run
  // right-preconditioned GMRES, which gives FGMRES's iterates when the
  // preconditioner is a fixed linear operator (every one here is); LGMRES is
  // rejected by the library with a message that names the key
  switch (config->getLinearSolverType()) {
    case LinearSolver::BICGSTAB: check(bes_implicit_set_solver(ctx, BES_SOLVER_BICGSTAB)); break;
    case LinearSolver::LGMRES: check(bes_implicit_set_solver(ctx, -1)); break;
    default: check(bes_implicit_set_solver(ctx, BES_SOLVER_GMRES)); break;
  }
  if (config->getPreconditioner() == 3) check(bes_lp_init(ctx, 20261202ull));
  if (config->getPreconditioner() == 4) {
    if (!config->getAmgAggregation())
      throw Error(BES_ERR_ARG, "AMG_PRECOND_COARSENER: only AGGREGATION is available");
    if (config->getAmgRelax() < 0)
      throw Error(BES_ERR_ARG, "AMG_PRECOND_SMOOTHER: DAMPED_JACOBI, SPAI0 or BLOCK_JACOBI");
    check(bes_amg_set_params(ctx, 0.08, 1.5, 0.72, 500, config->getAmgRelax(), 1));
  }
}

// ---------------------------------------------------------------------------
// CoefficientMatrix: host CSR in the reference layout, solved on the GPU
// ---------------------------------------------------------------------------
class CoefficientMatrix {
 public:
  // CoefficientMatrix.cpp:114-164: pattern from the mesh adjacency
  void initialize(Config *config, Mesh *mesh) {
    ctx_ = mesh->ctx();
    nv_ = config->getNumConservedVariable();
    check(bes_implicit_setup(ctx_, config->getPrecision(), config->getPreconditioner(), 30,
                             config->getLinearSolverNumIter(), config->getLinearSolverTol()));
    setupPreconditioner_(ctx_, config);
    const long nnz = bes_implicit_nnz(ctx_);
    const long nrow = (long)mesh->getNumElement() * nv_;
    row_.assign(nrow + 1, 0);
    col_.assign(nnz, 0);
    val_.assign(nnz, 0.0);
    // the device holds no system yet: build the pattern from the geometry
    int nd;
    long nc, nf;
    check(bes_ctx_info(ctx_, &nd, &nc, &nf));
    const int k = nd + 1;
    std::vector<long> adj(nc * k);
    check(bes_get_geometry(ctx_, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                           nullptr, adj.data(), nullptr, nullptr));
    long z = 0, r = 0;
    for (long e = 0; e < nc; e++)
      for (int a = 0; a < nv_; a++) {
        row_[r++] = z;
        for (int b = 0; b < nv_; b++) col_[z++] = e * nv_ + b;
        for (int m = 0; m < k; m++)
          if (adj[e * k + m] >= 0) for (int b = 0; b < nv_; b++) col_[z++] = adj[e * k + m] * nv_ + b;
      }
    row_[r] = z;
  }
  void changeAllVal(double x) { std::fill(val_.begin(), val_.end(), x); }
  void addToDiagonal(double x) { for (long i = 0; i + 1 < (long)row_.size(); i++) addToVal_(i, i, x); }
  void addToVal(unsigned long re, unsigned long ce, short rv, short cv, double x) {
    addToVal_(re * nv_ + rv, ce * nv_ + cv, x);
  }
  void addToBlock(unsigned long re, unsigned long ce, double multiplier, double *rowMajor) {
    for (int a = 0; a < nv_; a++)
      for (int b = 0; b < nv_; b++) addToVal_(re * nv_ + a, ce * nv_ + b, multiplier * rowMajor[a * nv_ + b]);
  }
  const std::vector<long> &rowPtr() const { return row_; }
  const std::vector<long> &colIdx() const { return col_; }
  std::vector<double> &values() { return val_; }
  // CoefficientMatrix.cpp:210-325: relative tolerance, returns iterations and error
  void solve(std::vector<double> &b, std::vector<double> &x, int &iter, double &error) {
    check(bes_implicit_set_system(ctx_, val_.data(), b.data()));
    int st = bes_linear_solve(ctx_, &iter, &error);
    if (st != BES_OK && st != BES_ERR_SOLVE) check(st, ctx_);
    x.resize(b.size());
    check(bes_get_linear_solution(ctx_, x.data()));
  }

 private:
  void addToVal_(long r, long c, double x) {
    for (long z = row_[r]; z < row_[r + 1]; z++)
      if (col_[z] == c) { val_[z] += x; return; }
  }
  bes_ctx *ctx_ = nullptr;
  int nv_ = 0;
  std::vector<long> row_, col_;
  std::vector<double> val_;
};

// ---------------------------------------------------------------------------
// Solvers (Solver.hpp:70-72): initialize, uniform initial condition, solve
// ---------------------------------------------------------------------------
class Solver {
 public:
  virtual ~Solver() {}
  virtual void initialize(Config *config, Gas *gas, Mesh *mesh) = 0;
  virtual void setUniversalInitialCondition(double *primitive) = 0;
  virtual void solve() = 0;
  // residual history: one row per pseudo-time step
  const std::vector<std::vector<double>> &history() const { return history_; }

 protected:
  void initializeBase_(Config *config, Gas *gas, Mesh *mesh) {
    config_ = config;
    gas_ = gas;
    mesh_ = mesh;
    check(bes_set_freestream(mesh->ctx(), config->getFreestreamPrimitiveVariables(0)));
  }
  // GasPerfect.tpp:181-216 for the uniform initial state
  void uniformState_(const double *P) {
    const int nd = mesh_->getNumDim(), nv = nd + 2;
    const double cv = config_->getGasConstant() / (config_->getRatioSpecificHeat() - 1.0);
    double U[5], q2 = 0;
    U[0] = P[0];
    for (int d = 0; d < nd; d++) { U[1 + d] = U[0] * P[1 + d]; q2 += P[1 + d] * P[1 + d]; }
    U[nv - 1] = U[0] * (cv * P[nd + 2] + q2 / 2.0);
    std::vector<double> all(mesh_->getNumElement() * nv);
    for (size_t e = 0; e < mesh_->getNumElement(); e++)
      for (int v = 0; v < nv; v++) all[e * nv + v] = U[v];
    check(bes_set_state(mesh_->ctx(), all.data()));
  }
  // Solver.cpp:123-128 with Solution.cpp:129-161: solution_<n>.vtk|csv and
  // slip_wall_<n>.csv after time step n (1-based), written every SAVE_INTERVAL
  // steps and after the last (SolverExplicit.cpp:425-433, SolverImplicit.cpp:588-595)
  void writeToFile_(unsigned long it) {
    const unsigned long n = config_->getNumTimeStep(), save = std::max(1ul, config_->getNumTimeStepBetweenSave());
    if (!(it % save == 0 || it == n - 1)) return;
    const std::string base = config_->getOutputFolderPath() + "/";
    const int fmt = config_->getOutputFormat();
    const std::string sol = base + "solution_" + std::to_string(it + 1) + (fmt == 1 ? ".csv" : ".vtk");
    check(bes_write_solution(mesh_->ctx(), sol.c_str(), fmt, 0));
    if (config_->getOutputSlipWall()) {
      const std::string sw = base + "slip_wall_" + std::to_string(it + 1) + ".csv";
      check(bes_write_slip_wall(mesh_->ctx(), sw.c_str(), 0));
    }
  }
  // Solver.cpp:130-137: residual.csv rows
  void appendToResidual_(const std::string &file, const std::string &row) {
    std::ofstream f(file, std::ios::app);
    f << row;
  }
  Config *config_ = nullptr;
  Gas *gas_ = nullptr;
  Mesh *mesh_ = nullptr;
  std::vector<std::vector<double>> history_;
};

class SolverExplicit : public Solver {
 public:
  void initialize(Config *config, Gas *gas, Mesh *mesh) override { initializeBase_(config, gas, mesh); }
  void setUniversalInitialCondition(double *primitive) override { uniformState_(primitive); }
  // SolverExplicit.cpp:395-438 (the freestream ramp is out of scope here)
  void solve() override {
    const std::string file = config_->getOutputFolderPath() + "/residual.csv";
    { std::ofstream f(file); }
    const int nv = config_->getNumConservedVariable();
    const unsigned long n = config_->getNumTimeStep();
    const std::vector<double> &alpha = config_->getTimeStepConstants();
    std::vector<double> linf(nv);
    for (unsigned long it = 0; it < n; it++) {
      check(bes_explicit_run(mesh_->ctx(), 1, (int)alpha.size(), alpha.data(), config_->getCourantNum(),
                             config_->getLocalTimeStepOption() == LocalTimeStep::YES, linf.data()),
            mesh_->ctx());
      std::ostringstream row;
      row << it + 1 << ",";
      for (int v = 0; v < nv; v++) { char b[32]; std::snprintf(b, sizeof b, "%e,", linf[v]); row << b; }
      row << "\n";
      appendToResidual_(file, row.str());
      history_.push_back(linf);
      writeToFile_(it);
    }
  }
};

class SolverImplicit : public Solver {
 public:
  void initialize(Config *config, Gas *gas, Mesh *mesh) override {
    initializeBase_(config, gas, mesh);
    check(bes_implicit_setup(mesh->ctx(), config->getPrecision(), config->getPreconditioner(), 30,
                             config->getLinearSolverNumIter(), config->getLinearSolverTol()));
    setupPreconditioner_(mesh->ctx(), config);
    check(bes_set_limiter_freeze(mesh->ctx(), config->getLimiterFreezeDrop()));
  }
  void setUniversalInitialCondition(double *primitive) override { uniformState_(primitive); }
  // SolverImplicit.cpp:552-600 with the adaptive CFL of :530-550
  void solve() override {
    const std::string file = config_->getOutputFolderPath() + "/residual.csv";
    { std::ofstream f(file); }
    const int nv = config_->getNumConservedVariable();
    const unsigned long n = config_->getNumTimeStep();
    const double cmax = config_->getAdaptiveCourantNumOption() == AdaptiveCourantNum::YES
                            ? config_->getAdaptiveCourantNumMax() : config_->getCourantNum();
    std::vector<double> h(n * (2 * nv + 3));
    int done = 0;
    check(bes_implicit_run(mesh_->ctx(), (int)n, config_->getCourantNum(), cmax,
                           config_->getAdaptiveCourantNumExponent(),
                           config_->getLocalTimeStepOption() == LocalTimeStep::YES, config_->getBlockRepeat(),
                           0.0, h.data(), &done),
          mesh_->ctx());
    for (int it = 0; it < done; it++) {
      const double *r = &h[it * (2 * nv + 3)];
      std::ostringstream row;
      row << it + 1 << ",";
      char b[32];
      for (int v = 0; v < nv; v++) { std::snprintf(b, sizeof b, "%e,", r[v]); row << b; }
      row << (int)r[2 * nv + 1] << ",";
      std::snprintf(b, sizeof b, "%e,", r[2 * nv + 2]);
      row << b << "\n";
      appendToResidual_(file, row.str());
      history_.push_back(std::vector<double>(r, r + 2 * nv + 3));
      writeToFile_(it);
    }
  }
};

}  // namespace bes_hip
