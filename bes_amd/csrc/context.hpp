// Device context of the BES solver core: geometry and state in HBM.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "flux.hpp"
#include "geometry.hpp"

namespace bes {

// Structure-of-arrays device mesh in locality (Morton) cell order.  Arrays
// indexed [m][...][N] hold one record per (cell, local face m); N = ncell.
struct DevMesh {
  int nd = 0, k = 0, N = 0;
  double *vol = nullptr;    // [N] cell volume (reference definition)
  double *eps2 = nullptr;   // [N] limiter epsilon^2 = ((K V^(1/nd))^(3/2))^2
  double *sproj = nullptr;  // [nd][N] projected lengths
  int *nbr = nullptr;       // [k][N] neighbour cell, -1 slip wall, -2 far field
  double *Sg = nullptr;     // [k][nd][N] area vector, outward from this cell
  double *w = nullptr;      // [k][N] inverse-distance weight of this cell
  double *rl = nullptr;     // [k][nd][N] centroid -> face centre
  double *nf = nullptr;     // [k][nd][N] face unit normal, owner -> neighbour
  double *af = nullptr;     // [k][N] face area
  double *rs = nullptr;     // [k][nd][N] reconstruction vector of this cell
  double *rn = nullptr;     // [k][nd][N] reconstruction vector of the neighbour
  uint32_t *info = nullptr; // [N] bits 0-7: face order (2 bits each); 8-11: owner flags
};

// Matrix storage of the implicit path (block CSR, own block first in each row).
struct DevBsr {
  int nnzb = 0;
  int *rowptr = nullptr;    // [N+1]
  int *col = nullptr;       // [nnzb]
  int *slot = nullptr;      // [k][N] block index of (cell, local face) neighbour, -1 none
  int *diag = nullptr;      // [N] block index of the diagonal
};

}  // namespace bes

struct bes_ctx {
  bes::Geometry geo;                 // reference-order geometry (host)
  std::vector<int64_t> perm, iperm;  // perm[new] = old cell, iperm[old] = new
  int nd = 0, nv = 0, k = 0, N = 0;
  bes::GasConst gc{};
  double venkat_k = 5.0;
  double finf[6] = {0, 0, 0, 0, 0, 0};
  bes::DevMesh dm;
  hipStream_t stream = nullptr;
  double *U0 = nullptr, *Ua = nullptr, *Ub = nullptr;  // [nv][N]
  double *LG = nullptr;        // [nv][nd][N] limited gradient phi*grad
  double *G = nullptr;         // [nv][nd][N] gradient (optional output)
  double *phi = nullptr;       // [nv][N] limiter (optional output)
  double *dt = nullptr;        // [N]
  double *dtg = nullptr;       // [1] global dt
  bool dt_local = true;        // mode of the last time-step computation
  double *R = nullptr;         // [nv][N] residual (optional output)
  unsigned long long *red = nullptr;   // reduction scratch, 64 words
  unsigned long long *err = nullptr;   // first failing (cell*k + m), ~0 if none
  double *linf = nullptr;              // [cap][nv] per-iteration Linf
  int linf_cap = 0;
  // implicit path
  bes::DevBsr bsr;
  void *impl = nullptr;        // implicit solver workspace (implicit.hip)
  long last_err_face = -1;
  // per-kernel HIP-event timing on this context's stream (bes_set_profiling)
  static constexpr int kProfKinds = 16;
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  struct PendingEvent { int kind; double bytes; hipEvent_t a, b; };
  std::vector<PendingEvent> ev_pending;
  double prof_ms[kProfKinds] = {0};
  double prof_bytes[kProfKinds] = {0};
  long prof_n[kProfKinds] = {0};
};

namespace bes {
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);
#define BES_HIP(call)                                            \
  do {                                                          \
    hipError_t e_ = (call);                                     \
    if (e_ != hipSuccess) return bes::hip_fail(e_, #call);      \
  } while (0)
inline int grid_for(int n, int bs) { return (n + bs - 1) / bs; }
// kernel kinds for per-kernel timing
enum ProfKind { kProfTimestep = 0, kProfGradlim = 1, kProfResidual = 2, kProfJacobian = 3,
                kProfAssemble = 4, kProfSpmv = 5, kProfPrecond = 6, kProfKrylov = 7, kProfFactor = 8,
                kProfUpdate = 9 };
// bracket a launch with events when profiling is on
// bytes: the launch's algorithmic HBM bytes (every array it must touch, once)
struct ProfScope {
  bes_ctx *c;
  int kind;
  double bytes;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(bes_ctx *c_, int kind_, double bytes_ = 0.0);
  ~ProfScope();
};
int prof_collect(bes_ctx *c);
}  // namespace bes
