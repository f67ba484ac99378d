// Implicit-path workspace and the block kernels shared by implicit.hip and
// learned.hip.  Matrix layout ("ELL-BSR"): slot 0 = diagonal block, slot 1+m =
// the block of the neighbour through local face m; entries stored [slot][e][N].
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <vector>

#include "context.hpp"
#include "kernels_common.hpp"

namespace bes {

constexpr int kRedBlocks = 1024;   // partial-sum blocks of the reductions

struct FaceDev {          // faces in device order (the flux pass's), device cell ids
  int F = 0;
  int *fo = nullptr;       // [F] owner
  int *fn = nullptr;       // [F] neighbour, -1 slip wall, -2 far field
  double *fnrm = nullptr;  // [nd][F] unit normal
  int *cf = nullptr;       // [k][N] device face id of (cell, local face)
  int *mrev = nullptr;     // [k][N] local index of this face in the neighbour
};

template <typename TM>
struct Mat {
  TM *val = nullptr;       // [(k+1)*nv*nv][N]
};

struct Impl {
  int prec = 64;           // 64 or 32
  int precond = 2;         // 0 none, 1 block Jacobi, 2 multicolour ILU(0), 3 learned, 4 AMG
  int restart = 30, maxiter = 1000;
  int solver = 0;          // 0 GMRES(restart) (also the reference's FGMRES), 1 BiCGStab
  double rtol = 1e-7;
  int repeat_nv = 1;
  FaceDev fd;
  void *JF = nullptr;      // [F][2][JP] face Jacobian pairs (TM), one padded row-major record per
                           // block, device face order
  double *b = nullptr;     // [N][nv] FP64 right-hand side
  double *x = nullptr;     // [N][nv] FP64 solution
  void *A = nullptr;       // matrix values (TM)
  void *Ac = nullptr;      // A in colour order, for the ILU(0) solves (TM)
  double *bc = nullptr;    // right-hand side in colour order
  double *xc = nullptr;    // solution in colour order
  void *LU = nullptr;      // factor / inverse diagonal blocks (TM), [(k+1)*nv*nv][N]
  void *Dr = nullptr;      // ILU(0): inverse diagonal blocks again, row-major [N][nv*nv] (TM)
  void *Ur = nullptr;      // ILU(0): upper blocks of each row, row-major [N][k][nv*nv] by position (TM)
  void *LM = nullptr;      // learned operator blocks (TM), [(k+1)*nv*nv][N] (learned.hip)
  __half *LMh = nullptr;   // FP16 copy applied by the FP32 solves (the operator's precision
                           // sets the convergence rate only, never the solution)
  void *lp = nullptr;      // learned-preconditioner network state (learned.hip)
  int lp_fallbacks = 0;    // solves that fell back to ILU(0)
  int lp_rebuild = 1;      // learned operator rebuilt at every lp_rebuild-th solve after an assembly
  int lp_since = 0;        // solves since its last build
  bool lp_current = false; // the stored operator belongs to the present parameters
  long lp_builds = 0;
  // online training: Adam steps on each assembled system before its solve
  int lp_online_steps = 0, lp_online_nprobe = 4, lp_online_nrich = 3;
  double lp_online_lr = 3e-3;
  unsigned long long lp_online_seed = 1;
  double lp_online_loss = 0;   // loss of the last online step
  void *amg = nullptr;     // aggregation AMG hierarchy (amg.hip)
  void *V = nullptr;       // Krylov basis (TV) [(m+1)][vstride]
  void *Z = nullptr;       // work vectors (TV) [3][vstride]
  size_t vstride = 0;      // N*nv rounded up to 64 elements
  double reorth_eta = 0.7; // DGKS re-orthogonalisation when ||w|| drops below eta of its norm
  long reorths = 0;        // re-orthogonalisations so far
  long extra_restarts = 0; // GMRES cycles whose Givens estimate met the target but the true residual did not
  double *part = nullptr;  // reduction partials [32][kRedBlocks] (BiCGStab)
  int lanes = 0;           // lanes L of the deterministic batched reductions
  double *S = nullptr;     // their lane sums [9][B][L]
  double *S2 = nullptr;    // level-1 sums of groups of 1024 lanes [9][B][L/1024]
  double *dots = nullptr;  // [64][B] reduced dot products, condition fastest
  double *hdots = nullptr; // [64][B] pinned host copy of dots (Arnoldi step)
  hipEvent_t ev_dots = nullptr;  // recorded after the copy into hdots
  unsigned long long *bits = nullptr;   // [2][8][B] Linf monitors of b and sum F*A, per variable and condition
  int ncolor = 0;
  int coloring = 0;               // greedy visiting order of the ILU(0) colouring (bes_implicit_set_coloring)
  std::vector<int> color_start;   // host: rows of colour c are crow[color_start[c] .. color_start[c+1])
  int *crow = nullptr;            // [N] rows grouped by colour: crow[p] = cell at position p
  int *color = nullptr;           // [N]
  // ILU(0) factors live in colour-major order: LU[slot][e][p]; per position p
  int *cpos = nullptr;            // [N] position of cell i
  int *cnb = nullptr;             // [k][N] neighbour cell through local face m of the cell at p (-1 none)
  int *cnp = nullptr;             // [k][N] that neighbour's position
  unsigned char *clow = nullptr;  // [N] bit m: neighbour m precedes in the elimination order
  long ninternal = 0;              // internal faces (off-diagonal block pairs)
  int last_iters = 0;
  double last_relres = 0;
  std::vector<int> iters_b;       // per condition: Krylov iterations of the last solve
  std::vector<double> relres_b;   // and its true relative residual
  std::vector<char> skip_b;       // conditions the next solve leaves alone (x = 0)
  bool assembled = false, factored = false;
  // pseudo-time loop state (bes_implicit_run / bes_implicit_continue)
  struct Run {
    bool active = false, reached = false;   // reached: every condition reached its drop
    double cfl_min = 0, cfl_max = 0, ex[2] = {1, 1}, stop_drop = 0;
    int local = 1, repeat_nv = 1;
    std::vector<double> cfl, r0;      // [B] Courant number of the next step, first density residual
    std::vector<double> lb_old;       // [B][nv] Linf of b of the previous step
    std::vector<char> done;           // [B] reached its drop (state frozen from then on)
    std::vector<int> steps;           // [B] steps taken
  } run;
};

// aggregation AMG (amg.hip): hierarchy from Impl::A, one V-cycle z = M r
int amg_build(bes_ctx *c, Impl *im);
template <typename TV>
int amg_apply(bes_ctx *c, Impl *im, const TV *r, TV *z, TV *tmp);
void amg_release(Impl *im);

// ---------------------------------------------------------------------------
// small dense block helpers (NV x NV, row-major, registers)
// ---------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_inverse(double *A, double *Inv) {
  // Gauss-Jordan with partial pivoting
#pragma unroll
  for (int r = 0; r < NV; r++)
#pragma unroll
    for (int c = 0; c < NV; c++) Inv[r * NV + c] = (r == c) ? 1.0 : 0.0;
#pragma unroll
  for (int p = 0; p < NV; p++) {
    int piv = p;
    double best = ::fabs(A[p * NV + p]);
#pragma unroll
    for (int r = p + 1; r < NV; r++) {
      double a = ::fabs(A[r * NV + p]);
      if (a > best) { best = a; piv = r; }
    }
    // rows p and piv swap; written as a conditional exchange with every row
    // below p so that all indices are compile-time constants and A and Inv
    // stay in registers (an exchange through a run-time row index put both
    // arrays in per-lane scratch memory for nv = 5)
#pragma unroll
    for (int r = p + 1; r < NV; r++) {
      const bool sw = piv == r;
#pragma unroll
      for (int c = 0; c < NV; c++) {
        const double a = A[p * NV + c], b = A[r * NV + c];
        A[p * NV + c] = sw ? b : a;
        A[r * NV + c] = sw ? a : b;
        const double x = Inv[p * NV + c], y = Inv[r * NV + c];
        Inv[p * NV + c] = sw ? y : x;
        Inv[r * NV + c] = sw ? x : y;
      }
    }
    const double inv = 1.0 / A[p * NV + p];
#pragma unroll
    for (int c = 0; c < NV; c++) { A[p * NV + c] *= inv; Inv[p * NV + c] *= inv; }
#pragma unroll
    for (int r = 0; r < NV; r++) {
      if (r == p) continue;
      const double f = A[r * NV + p];
#pragma unroll
      for (int c = 0; c < NV; c++) {
        A[r * NV + c] -= f * A[p * NV + c];
        Inv[r * NV + c] -= f * Inv[p * NV + c];
      }
    }
  }
}

template <typename TM, int NB>
__device__ __forceinline__ void load_block(const TM *__restrict__ base, int slot, int N, int i, double *B) {
#pragma unroll
  for (int e = 0; e < NB; e++) B[e] = (double)base[((size_t)slot * NB + e) * N + i];
}
template <typename TM, int NB>
__device__ __forceinline__ void store_block(TM *__restrict__ base, int slot, int N, int i, const double *B) {
#pragma unroll
  for (int e = 0; e < NB; e++) base[((size_t)slot * NB + e) * N + i] = (TM)B[e];
}
// the same for a batch of conditions (condition b fastest): entry e of slot s
// of row i at ((s*NB + e)*N + i)*nb + b; nb = 1 gives the layouts above
// (addressed from the block's first entry with a uniform stride: a full index
// per entry costs a 64-bit register pair per access in flight)
template <typename TM, int NB>
__device__ __forceinline__ void load_block_b(const TM *__restrict__ base, int slot, int N, int i, int nb, int b,
                                             double *B) {
  const TM *q = base + ((size_t)slot * NB * N + i) * nb + b;
  const size_t es = (size_t)N * nb;
#pragma unroll
  for (int e = 0; e < NB; e++) B[e] = (double)q[e * es];
}
template <typename TM, int NB>
__device__ __forceinline__ void store_block_b(TM *__restrict__ base, int slot, int N, int i, int nb, int b,
                                              const double *B) {
  TM *q = base + ((size_t)slot * NB * N + i) * nb + b;
  const size_t es = (size_t)N * nb;
#pragma unroll
  for (int e = 0; e < NB; e++) q[e * es] = (TM)B[e];
}

__device__ __forceinline__ double to_d(double x) { return x; }
__device__ __forceinline__ double to_d(float x) { return (double)x; }
__device__ __forceinline__ double to_d(__half x) { return (double)__half2float(x); }

// ---------------------------------------------------------------------------
// y = A x for an ELL-BSR matrix (vectors [N][nv], type TV)
// ---------------------------------------------------------------------------
// y = A x for B = 1 << lgB conditions at once: matrix entries at
// ((slot*NB + e)*N + i)*B + b, vectors at (i*nv + v)*B + b (condition fastest)
template <int ND, typename TM, typename TV>
__global__ __launch_bounds__(256) void k_spmv(const TM *__restrict__ A, DevMesh dm, const TV *__restrict__ x,
                                              TV *__restrict__ y, int lgB) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int N = dm.N, B = 1 << lgB;
  const int t = cell_index();
  const int i = t >> lgB, b = t & (B - 1);
  if (i >= N) return;
  // block entries from the row's first entry with the uniform stride es
  const size_t es = (size_t)N * B;
  const TM *Ar = A + (size_t)i * B + b;
  double acc[NV];
  double xi[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) xi[c] = (double)x[((size_t)i * NV + c) * B + b];
#pragma unroll
  for (int r = 0; r < NV; r++) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < NV; c++) s += to_d(Ar[(r * NV + c) * es]) * xi[c];
    acc[r] = s;
  }
#pragma unroll
  for (int m = 0; m < K; m++) {
    const int j = dm.nbr[(size_t)m * N + i];
    if (j < 0) continue;
    double xj[NV];
#pragma unroll
    for (int c = 0; c < NV; c++) xj[c] = (double)x[((size_t)j * NV + c) * B + b];
    const TM *Ao = Ar + (size_t)(1 + m) * NB * es;
#pragma unroll
    for (int r = 0; r < NV; r++) {
      double s = 0;
#pragma unroll
      for (int c = 0; c < NV; c++) s += to_d(Ao[(r * NV + c) * es]) * xj[c];
      acc[r] += s;
    }
  }
#pragma unroll
  for (int r = 0; r < NV; r++) y[((size_t)i * NV + r) * B + b] = (TV)acc[r];
}

}  // namespace bes
