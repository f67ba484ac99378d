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
  int *cfd = nullptr;       // [k][N] device face index of (cell, local face)
  uint32_t *info = nullptr; // [N] bits 0-7: face order (2 bits each); 8-11: owner flags
  int *cfq = nullptr;       // [k][N] device face id of the cell's q-th face in reference face order
  double *saq = nullptr;    // [k][N] that face's area, negated where the cell is not the owner
  // faces in device order: sorted by the owner's (Morton) index, then local face
  int F = 0;
  int *fo = nullptr;        // [F] owner cell
  int *fnb = nullptr;       // [F] neighbour cell, -1 slip wall, -2 far field
  double *fnr = nullptr;    // [nd][F] unit normal, owner -> neighbour
  double *frs = nullptr;    // [nd][F] owner's reconstruction vector for the face
  double *frn = nullptr;    // [nd][F] neighbour's reconstruction vector for the face
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
  std::vector<double> node_xyz;      // [nnode][nd] mesh nodes (solution files)
  std::vector<int64_t> cell_nodes;   // [ncell][nd+1] cell vertices (solution files)
  std::vector<int64_t> perm, iperm;  // perm[new] = old cell, iperm[old] = new
  int nd = 0, nv = 0, k = 0, N = 0;
  bes::GasConst gc{};
  double venkat_k = 5.0;
  // B freestream conditions share the mesh (B = 1 << lgB <= kMaxBatch); every
  // per-cell device array below carries the condition as its fastest index
  static constexpr int kMaxBatch = 64;
  int B = 1, lgB = 0;
  double finf[6] = {0, 0, 0, 0, 0, 0};  // freestream primitive of condition 0
  std::vector<double> finf_b;           // [B][nv+1] freestream primitive per condition
  double *dfinf = nullptr;              // device copy of finf_b
  bes::DevMesh dm;
  hipStream_t stream = nullptr;
  double *U0 = nullptr, *Ua = nullptr, *Ub = nullptr;  // [nv][N][B]
  double *LG = nullptr;        // [N][nv][1+nd][B] per-cell pack: per variable U_v, then phi*grad U_v [d]
  double *FF = nullptr;        // [F][fs][B] face fluxes per unit area (device face order)
  int fs = 0;                  // face-flux record length: nv, rounded up to even for B = 1 (16-byte access)
  std::vector<int64_t> fref;   // reference face id of each device face
  double *G = nullptr;         // [nv][nd][N][B] gradient (optional output)
  double *phi = nullptr;       // [nv][N][B] limiter (optional output)
  double *dt = nullptr;        // [N][B]
  double *dtg = nullptr;       // [B] global dt
  bool dt_local = true;        // mode of the last time-step computation
  double *R = nullptr;         // [nv][N][B] residual (optional output)
  // limiter freeze (bes_set_limiter_freeze): every gradient pass stores the
  // limiter of the conditions not yet frozen in phi_frz; a frozen condition's
  // passes read it back instead of recomputing it
  bool lim_freeze_on = false;        // phi_frz in use
  double lim_freeze_drop = 0.0;      // drop that freezes a condition in the implicit loop (0: none)
  double *phi_frz = nullptr;         // [nv][N][B]
  std::vector<char> lim_frozen;      // [B]
  int lim_epoch = 0;                 // changes with the freeze state (graph captures hold it)
  // reduction scratch: [0, 64) global-dt minimum per condition, [64, 64 + nv*B)
  // Linf bit patterns [v][B]
  static constexpr int kRedWords = 64 + 5 * kMaxBatch;
  unsigned long long *red = nullptr;
  unsigned long long *lbits = nullptr;  // Linf monitor words [kLinfSlots][nv][B] (see kernels_common.hpp)
  unsigned long long *err = nullptr;   // first failing device face, ~0 if none
  double *linf = nullptr;              // [cap][B][nv] per-iteration Linf
  int linf_cap = 0;
  // pseudo-time step captured as a HIP graph (bes_explicit_graph)
  bool use_graph = false;
  // all steps of a run in one cooperative launch (bes_explicit_persistent)
  bool use_persistent = false;
  hipGraphExec_t graph = nullptr;
  int graph_key[5] = {0, 0, 0, 0, 0};  // nstage, local, linf on, B, limiter-freeze epoch of the capture
  double graph_cfl = 0.0;
  std::vector<double> graph_alpha;
  int *dit = nullptr;                  // device iteration counter of graph replays
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
  // block size of each kernel kind (0: 256), set by bes_autotune
  int tune_bs[kProfKinds] = {0};
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
inline int tuned_bs(const bes_ctx *c, int kind) { return c->tune_bs[kind] > 0 ? c->tune_bs[kind] : 256; }
// kernel kinds for per-kernel timing
enum ProfKind { kProfTimestep = 0, kProfGradlim = 1, kProfResidual = 2, kProfJacobian = 3,
                kProfAssemble = 4, kProfSpmv = 5, kProfPrecond = 6, kProfKrylov = 7, kProfFactor = 8,
                kProfUpdate = 9, kProfFaceFlux = 10, kProfLearned = 11 };
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
