// Aggregation AMG preconditioner for the implicit systems (MI355X, gfx950).
//
// The reference preconditions its GMRES with an AMGCL hierarchy rebuilt at
// every nonlinear step (CoefficientMatrix.cpp:210-325: coarsening
// "aggregation" with aggr.block_size = nv, relaxation from the .besi).  This
// file builds and applies that kind of hierarchy on the GPU from the ELL-BSR
// matrix of the implicit path:
//
//   strength   one scalar per nv x nv block, its largest |entry|; block (i, j)
//              is strong when eps^2 d_i d_j < s_ij^2 (d the diagonal block's
//              scalar), and rows with no strong block are left out
//              (pointwise_aggregates.hpp:188-260, plain_aggregates.hpp:118-150)
//   aggregates roots pairwise at least three strong links apart, chosen in
//              rounds: an undecided row becomes a root when its hashed
//              priority is the largest among undecided rows within two links,
//              and every row within two links of a root is then decided.  A
//              row one link from a root joins it; the others join the
//              aggregate of their strongest neighbour one link from a root.
//              The reference sweeps the rows in order instead
//              (plain_aggregates.hpp:160-190); both give aggregates of radius
//              two around a root, this one in O(log n) parallel rounds.
//   transfer   piecewise-constant P per component, R = P^T
//              (tentative_prolongation.hpp:195-205)
//   coarse     A_c = P^T A P / over_interp as block sums in a fixed order
//              (aggregation.hpp:188-196); one wave per coarse row
//   coarsest   at most coarse_enough scalar rows: explicit inverse by
//              Gauss-Jordan with partial pivoting, applied as a dense product
//              (the reference factors it, default_direct_solver.hpp)
//   relax      damped Jacobi (scalar diagonal), SPAI0, or damped block Jacobi
//              (nv x nv diagonal-block inverses; no reference counterpart)
//   cycle      V-cycle from x = 0 with one pre- and one post-sweep
//              (amg.hpp:287-304, 498-541)
//
// Level 0 is the implicit path's matrix itself (device cell order, ELL-BSR);
// coarse levels are block CSR with ascending columns, blocks stored
// contiguously ([nnzb][nv*nv]).  Every sum runs in a fixed order, so the
// hierarchy and the cycle are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bes_hip.h"
#include "implicit.hpp"

namespace bes {
namespace amg {

enum Relax { kJacobi = 0, kSpai0 = 1, kBlockJacobi = 2 };
enum State : unsigned char { kUndecided = 0, kRoot = 1, kMember = 2, kRemoved = 3 };

constexpr int kMaxLevels = 20;
constexpr int kMaxRounds = 200;      // distance-2 independent-set rounds (O(log n) expected)
constexpr int kDenseMax = 1024;      // largest coarsest system inverted densely (scalar rows)

struct Params {
  double eps_strong = 0.08, over_interp = 1.5, damping = 0.72;
  int coarse_enough = 500;
  int relax = kJacobi;
  int rebuild = 1;          // hierarchy rebuilt at every rebuild-th solve after an assembly
  int sweep = 0;            // 1: AMGCL's serial row sweep instead of the parallel aggregation
};

struct Level {
  int n = 0;                // block rows
  long nnzb = 0;            // stored blocks
  // coarse levels: block CSR, ascending columns, diagonal position
  int *rp = nullptr, *col = nullptr, *dpos = nullptr;
  void *val = nullptr;      // [nnzb][nv*nv] (matrix type)
  void *w = nullptr;        // relaxation (matrix type): [n][nv] scalar factors or [n][nv*nv] damped inverse blocks
  void *f = nullptr, *u = nullptr, *t = nullptr, *xa = nullptr;   // vectors [n][nv] (vector type)
  // coarsening to the next level
  int *agg = nullptr;       // [n] aggregate (= next-level row), -1 removed
  unsigned char *state = nullptr;   // [n] final State of the row
  int *mptr = nullptr, *mem = nullptr;   // members of next-level row I: mem[mptr[I] .. mptr[I+1]), ascending
  // FP32 solves, coarse levels: row-scaled FP16 copy of val for the sweeps
  __half *hval = nullptr;
  float *hrm = nullptr;
};

struct Amg {
  Params prm;
  int built_type = 0;       // sizeof(TM) of the built hierarchy, 0 none
  long builds = 0;          // hierarchies built so far
  int reused = 0;           // solves since the last build
  bool force = true;        // next request builds (new parameters)
  std::vector<Level> lv;
  int nlev = 0;
  bool direct = false;
  int ndense = 0;           // scalar rows of the coarsest system when direct
  double *dense = nullptr, *dense2 = nullptr, *inv = nullptr;
  // scratch sized for level 0
  double *s = nullptr, *dia = nullptr;
  unsigned char *strong = nullptr, *sym = nullptr, *near = nullptr;
  unsigned long long *m1 = nullptr;
  int *iflag = nullptr, *ibuf = nullptr, *ibuf2 = nullptr;
  int *flags = nullptr;     // device: [0] undecided count, [1] error bits, [2] pivot failures
  std::vector<int *> scan_tmp;
  void *t0 = nullptr;       // level-0 residual (vector type)
  // FP32 solves: row-scaled FP16 copies of the level-0 matrix and relaxation
  // data for the cycle (the hierarchy is built from the full-precision matrix)
  __half *Ah = nullptr, *Wh = nullptr;
  float *rm = nullptr;      // [N] row scale m_i: A_i = m_i Ah_i, W_i = Wh_i / m_i
  std::vector<void *> owned;     // allocations outside the arena
  void *arena = nullptr;
  size_t arena_cap = 0, arena_off = 0, arena_need = 0;
};

static Amg *amg_of(Impl *im) { return (Amg *)im->amg; }

// ---------------------------------------------------------------------------
// row access: level 0 (ELL-BSR of the implicit path) and coarse levels (CSR)
// ---------------------------------------------------------------------------
template <typename TM, int NV>
struct EllAcc {
  static constexpr int NB = NV * NV;
  const TM *A;
  const int *nbr, *mrev;
  int N, K;
  __device__ int len(int) const { return K + 1; }
  __device__ int col(int i, int q) const { return q == 0 ? i : nbr[(size_t)(q - 1) * N + i]; }
  __device__ size_t eid(int i, int q) const { return (size_t)q * N + i; }
  __device__ int diag(int) const { return 0; }
  __device__ double v(int i, int q, int e) const { return to_d(A[((size_t)q * NB + e) * N + i]); }
  // packed (row, entry) for the Galerkin gather; rows < 2^28
  __device__ int code(int i, int q) const { return (i << 3) | q; }
  __device__ double vc(int cd, int e) const { return v(cd >> 3, cd & 7, e); }
  // entry of row j (a neighbour through entry q of row i) that points back to i
  __device__ int back(int i, int q, int) const { return 1 + mrev[(size_t)(q - 1) * N + i]; }
};

// level 0 from the row-scaled FP16 copy (cycle only)
template <int NV>
struct EllAccH {
  static constexpr int NB = NV * NV;
  const __half *A;
  const float *rm;
  const int *nbr;
  int N, K;
  __device__ int len(int) const { return K + 1; }
  __device__ int col(int i, int q) const { return q == 0 ? i : nbr[(size_t)(q - 1) * N + i]; }
  __device__ double v(int i, int q, int e) const {
    return (double)__half2float(A[((size_t)q * NB + e) * N + i]) * (double)rm[i];
  }
};

template <typename TM, int NV>
struct CsrAcc {
  static constexpr int NB = NV * NV;
  const int *rp, *cl, *dp;
  const TM *val;
  __device__ int len(int i) const { return rp[i + 1] - rp[i]; }
  __device__ int col(int i, int q) const { return cl[rp[i] + q]; }
  __device__ size_t eid(int i, int q) const { return (size_t)rp[i] + q; }
  __device__ int diag(int i) const { return dp[i]; }
  __device__ double v(int i, int q, int e) const { return to_d(val[((size_t)rp[i] + q) * NB + e]); }
  __device__ int code(int i, int q) const { return rp[i] + q; }
  __device__ double vc(int cd, int e) const { return to_d(val[(size_t)cd * NB + e]); }
  // entry of row j that holds column i (columns ascending), -1 if none
  __device__ int find(int j, int i) const {
    int lo = rp[j], hi = rp[j + 1] - 1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1, cm = cl[mid];
      if (cm == i) return mid - rp[j];
      if (cm < i) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
  }
};

template <typename TM, int NV>
__device__ __forceinline__ int back_entry(const EllAcc<TM, NV> &a, int i, int q, int j) { return a.back(i, q, j); }
template <typename TM, int NV>
__device__ __forceinline__ int back_entry(const CsrAcc<TM, NV> &a, int i, int, int j) { return a.find(j, i); }

__device__ __forceinline__ unsigned hash32(unsigned x) {
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
  return x;
}
__device__ __forceinline__ unsigned long long prio(int i, unsigned salt) {
  return ((unsigned long long)(hash32((unsigned)i ^ salt) | 1u) << 32) | (unsigned)i;
}

__device__ __forceinline__ int row_index() { return blockIdx.x * blockDim.x + threadIdx.x; }

// ---------------------------------------------------------------------------
// strength and aggregation kernels (one thread per row)
// ---------------------------------------------------------------------------
template <int NV, typename Acc>
__global__ __launch_bounds__(256) void k_pointwise(Acc a, int n, double *__restrict__ s, double *__restrict__ dia) {
  const int i = row_index();
  if (i >= n) return;
  const int L = a.len(i), dq = a.diag(i);
  for (int q = 0; q < L; q++) {
    const int j = a.col(i, q);
    double m = 0;
    if (j >= 0)
      for (int e = 0; e < NV * NV; e++) m = fmax(m, fabs(a.v(i, q, e)));
    s[a.eid(i, q)] = m;
    if (q == dq) dia[i] = m;
  }
}

template <typename Acc>
__global__ __launch_bounds__(256) void k_strong(Acc a, int n, const double *__restrict__ s,
                                                const double *__restrict__ dia, double e2,
                                                unsigned char *__restrict__ strong, unsigned char *__restrict__ state) {
  const int i = row_index();
  if (i >= n) return;
  const int L = a.len(i), dq = a.diag(i);
  const double edi = e2 * dia[i];
  bool any = false;
  for (int q = 0; q < L; q++) {
    const int j = a.col(i, q);
    const double v = s[a.eid(i, q)];
    const bool st = j >= 0 && q != dq && edi * dia[j] < v * v;
    strong[a.eid(i, q)] = st;
    any |= st;
  }
  state[i] = any ? kUndecided : kRemoved;
}

// symmetric strength among kept rows: strong(i -> j) or strong(j -> i)
template <typename Acc>
__global__ __launch_bounds__(256) void k_sym(Acc a, int n, const unsigned char *__restrict__ strong,
                                             const unsigned char *__restrict__ state, unsigned char *__restrict__ sym) {
  const int i = row_index();
  if (i >= n) return;
  const int L = a.len(i), dq = a.diag(i);
  const bool keep = state[i] != kRemoved;
  for (int q = 0; q < L; q++) {
    const int j = a.col(i, q);
    bool sy = false;
    if (keep && j >= 0 && q != dq && state[j] != kRemoved) {
      sy = strong[a.eid(i, q)];
      if (!sy) {
        const int b = back_entry(a, i, q, j);
        sy = b >= 0 && strong[a.eid(j, b)];
      }
    }
    sym[a.eid(i, q)] = sy;
  }
}

template <typename Acc>
__global__ __launch_bounds__(256) void k_mis_max1(Acc a, int n, const unsigned char *__restrict__ state,
                                                  const unsigned char *__restrict__ sym, unsigned salt,
                                                  unsigned long long *__restrict__ m1) {
  const int i = row_index();
  if (i >= n) return;
  unsigned long long m = state[i] == kUndecided ? prio(i, salt) : 0ull;
  const int L = a.len(i);
  for (int q = 0; q < L; q++) {
    if (!sym[a.eid(i, q)]) continue;
    const int j = a.col(i, q);
    if (state[j] == kUndecided) m = max(m, prio(j, salt));
  }
  m1[i] = m;
}

template <typename Acc>
__global__ __launch_bounds__(256) void k_mis_max2(Acc a, int n, unsigned char *__restrict__ state,
                                                  const unsigned char *__restrict__ sym, unsigned salt,
                                                  const unsigned long long *__restrict__ m1) {
  const int i = row_index();
  if (i >= n || state[i] != kUndecided) return;
  unsigned long long m = m1[i];
  const int L = a.len(i);
  for (int q = 0; q < L; q++)
    if (sym[a.eid(i, q)]) m = max(m, m1[a.col(i, q)]);
  if (m == prio(i, salt)) state[i] = kRoot;   // writes only its own row; others read m1
}

template <typename Acc>
__global__ __launch_bounds__(256) void k_mis_near(Acc a, int n, const unsigned char *__restrict__ state,
                                                  const unsigned char *__restrict__ sym,
                                                  unsigned char *__restrict__ near) {
  const int i = row_index();
  if (i >= n) return;
  bool r = state[i] == kRoot;
  const int L = a.len(i);
  for (int q = 0; q < L && !r; q++)
    if (sym[a.eid(i, q)] && state[a.col(i, q)] == kRoot) r = true;
  near[i] = r;
}

template <typename Acc>
__global__ __launch_bounds__(256) void k_mis_cover(Acc a, int n, unsigned char *__restrict__ state,
                                                   const unsigned char *__restrict__ sym,
                                                   const unsigned char *__restrict__ near, int *__restrict__ undecided) {
  const int i = row_index();
  if (i >= n || state[i] != kUndecided) return;
  bool c = near[i];
  const int L = a.len(i);
  for (int q = 0; q < L && !c; q++)
    if (sym[a.eid(i, q)] && near[a.col(i, q)]) c = true;
  if (c) state[i] = kMember;
  else atomicAdd(undecided, 1);
}

__global__ void k_root_flag(int n, const unsigned char *__restrict__ state, int *__restrict__ flag) {
  const int i = row_index();
  if (i < n) flag[i] = state[i] == kRoot;
}

// a row one link from a root joins it (the root is unique: roots are three links apart)
template <typename Acc>
__global__ __launch_bounds__(256) void k_assign1(Acc a, int n, const unsigned char *__restrict__ state,
                                                 const unsigned char *__restrict__ sym, const int *__restrict__ rid,
                                                 int *__restrict__ agg) {
  const int i = row_index();
  if (i >= n) return;
  const unsigned char st = state[i];
  int g = -2;
  if (st == kRoot) g = rid[i];
  else if (st == kRemoved) g = -1;
  else {
    const int L = a.len(i);
    for (int q = 0; q < L; q++)
      if (sym[a.eid(i, q)]) {
        const int j = a.col(i, q);
        if (state[j] == kRoot) { g = rid[j]; break; }
      }
  }
  agg[i] = g;
}

// the other rows (two links from a root) join their strongest assigned neighbour
template <typename Acc>
__global__ __launch_bounds__(256) void k_assign2(Acc a, int n, const unsigned char *__restrict__ sym,
                                                 const double *__restrict__ s, const int *__restrict__ ain,
                                                 int *__restrict__ aout, int *__restrict__ err) {
  const int i = row_index();
  if (i >= n) return;
  int g = ain[i];
  if (g == -2) {
    double best = -1.0;
    int bj = 0x7fffffff;
    const int L = a.len(i);
    for (int q = 0; q < L; q++) {
      if (!sym[a.eid(i, q)]) continue;
      const int j = a.col(i, q);
      const int gj = ain[j];
      if (gj < 0) continue;
      const double v = s[a.eid(i, q)];
      if (v > best || (v == best && j < bj)) { best = v; bj = j; g = gj; }
    }
    if (g < 0) { atomicOr(err, 1); g = -1; }
  }
  aout[i] = g;
}

__global__ void k_fill_i(int *p, long n, int v) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_count(int n, const int *__restrict__ agg, int *__restrict__ cnt) {
  const int i = row_index();
  if (i < n && agg[i] >= 0) atomicAdd(&cnt[agg[i]], 1);
}

__global__ void k_scatter(int n, const int *__restrict__ agg, const int *__restrict__ mptr, int *__restrict__ fill,
                          int *__restrict__ mem) {
  const int i = row_index();
  if (i >= n || agg[i] < 0) return;
  const int g = agg[i];
  mem[mptr[g] + atomicAdd(&fill[g], 1)] = i;
}

__global__ void k_sort_members(int nc, const int *__restrict__ mptr, int *__restrict__ mem) {
  const int I = row_index();
  if (I >= nc) return;
  int *m = mem + mptr[I];
  const int c = mptr[I + 1] - mptr[I];
  for (int x = 1; x < c; x++) {
    const int v = m[x];
    int y = x - 1;
    while (y >= 0 && m[y] > v) { m[y + 1] = m[y]; y--; }
    m[y + 1] = v;
  }
}

// AMGCL's row sweep (plain_aggregates.hpp:137-206) in one thread, for checking
// the hierarchy against the reference's own aggregates: rows in the order
// order[r] (the reference's cell order on level 0; row order on coarse levels),
// each unassigned row a root that takes its strong neighbours (even taken
// ones) and, provisionally, their unassigned strong neighbours; emptied
// aggregates are dropped.  Serial, so meant for verification-size meshes.
template <typename Acc>
__global__ void k_row_sweep(Acc a, int n, const int *__restrict__ order, const unsigned char *__restrict__ strong,
                            int *__restrict__ id, unsigned char *__restrict__ state, int *__restrict__ used,
                            int *__restrict__ naggr) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int UNDEF = -1, REM = -2;
  for (int i = 0; i < n; i++) id[i] = state[i] == kRemoved ? REM : UNDEF;
  int count = 0;
  for (int r = 0; r < n; r++) {
    const int i = order ? order[r] : r;
    if (id[i] != UNDEF) continue;
    const int cur = count++;
    id[i] = cur;
    state[i] = kRoot;
    const int L = a.len(i);
    for (int q = 0; q < L; q++) {
      if (!strong[a.eid(i, q)]) continue;
      const int c = a.col(i, q);
      if (id[c] != REM) {
        id[c] = cur;
        state[c] = kMember;
      }
    }
    for (int q = 0; q < L; q++) {   // the neighbours just taken: strong, not removed
      if (!strong[a.eid(i, q)]) continue;
      const int c = a.col(i, q);
      if (id[c] == REM) continue;
      const int Lc = a.len(c);
      for (int qq = 0; qq < Lc; qq++) {
        if (!strong[a.eid(c, qq)]) continue;
        const int cc = a.col(c, qq);
        if (id[cc] == UNDEF) {
          id[cc] = cur;
          state[cc] = kMember;
        }
      }
    }
  }
  for (int g = 0; g < count; g++) used[g] = 0;
  for (int i = 0; i < n; i++)
    if (id[i] >= 0) used[id[i]] = 1;
  int m = 0;
  for (int g = 0; g < count; g++) {
    const int u = used[g];
    used[g] = m;
    m += u;
  }
  for (int i = 0; i < n; i++) id[i] = id[i] >= 0 ? used[id[i]] : -1;
  *naggr = m;
}

// ---------------------------------------------------------------------------
// exclusive scan of int arrays (block scan, scan of block sums, add back)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_scan_block(const int *__restrict__ in, int *__restrict__ out, int n,
                                                     int *__restrict__ bsum) {
  __shared__ int sh[1024];
  const int t = threadIdx.x, i = blockIdx.x * 1024 + t;
  const int v = i < n ? in[i] : 0;
  sh[t] = v;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int add = t >= o ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  if (i < n) out[i] = sh[t] - v;
  if (t == 1023) bsum[blockIdx.x] = sh[1023];
}

__global__ void k_scan_add(int *__restrict__ out, int n, const int *__restrict__ boff) {
  const int i = blockIdx.x * 1024 + threadIdx.x;
  if (i < n) out[i] += boff[blockIdx.x];
}

// ---------------------------------------------------------------------------
// Galerkin product, one wave per coarse row
// ---------------------------------------------------------------------------
template <int HASH>
__device__ __forceinline__ int ht_insert(int *ht, int key, int *count) {
  unsigned h = ((unsigned)key * 2654435761u) & (HASH - 1);
  for (int probe = 0; probe < HASH; probe++) {
    const int old = atomicCAS(&ht[h], -1, key);
    if (old == -1) { atomicAdd(count, 1); return 1; }
    if (old == key) return 0;
    h = (h + 1) & (HASH - 1);
  }
  return -1;
}

// LDS sizes of the Galerkin kernels: a small configuration for rows with few
// gathered blocks (most levels; high occupancy) and a large fallback
struct GalSmall { static constexpr int HASH = 256, COLS = 128, CAND = 256, THREADS = 64; };
struct GalMedium { static constexpr int HASH = 512, COLS = 256, CAND = 1024, THREADS = 128; };
struct GalLarge { static constexpr int HASH = 2048, COLS = 1024, CAND = 2048, THREADS = 256; };

// gathers the coarse columns of row I into the LDS table (one block per coarse
// row); returns false when the row does not fit the configuration
template <typename G, typename Acc>
__device__ bool gather_cols(const Acc &a, int I, const int *__restrict__ mptr, const int *__restrict__ mem,
                            const int *__restrict__ agg, int *ht, int *cnt, int *bad) {
  const int t = threadIdx.x;
  for (int h = t; h < G::HASH; h += G::THREADS) ht[h] = -1;
  if (t == 0) { *cnt = 0; *bad = 0; }
  __syncthreads();
  const int m0 = mptr[I], m1 = mptr[I + 1];
  int off = 0;
  for (int mi = m0; mi < m1; mi++) {
    const int i = mem[mi];
    const int L = a.len(i);
    if (off + L <= G::CAND)
      for (int q = t; q < L; q += G::THREADS) {
        const int j = a.col(i, q);
        const int J = j >= 0 ? agg[j] : -1;
        if (J >= 0 && ht_insert<G::HASH>(ht, J, cnt) < 0) atomicOr(bad, 1);
      }
    off += L;
  }
  __syncthreads();
  return !(*bad) && *cnt <= G::COLS && off <= G::CAND && (*cnt) * 2 <= G::HASH;
}

// unique coarse columns per coarse row; err |= fail_bit when a row does not fit
template <typename G, typename Acc>
__global__ __launch_bounds__(G::THREADS) void k_gal_count(Acc a, int nc, const int *__restrict__ mptr,
                                                          const int *__restrict__ mem, const int *__restrict__ agg,
                                                          int *__restrict__ rowcnt, int *__restrict__ err,
                                                          int fail_bit) {
  __shared__ int ht[G::HASH];
  __shared__ int cnt, bad;
  const int I = blockIdx.x;
  if (I >= nc) return;
  const bool ok = gather_cols<G>(a, I, mptr, mem, agg, ht, &cnt, &bad);
  if (threadIdx.x == 0) {
    rowcnt[I] = cnt;
    if (!ok) atomicOr(err, fail_bit);
  }
}

// Coarse row I: columns sorted ascending; the gathered fine blocks (members in
// ascending order, entries in row order) are placed by a stable counting sort
// on their coarse column, so that block (I, J) is the sum of its fine blocks
// in that order.
template <typename G, int NV, typename Acc, typename TMc>
__global__ __launch_bounds__(G::THREADS) void k_gal_fill(Acc a, int nc, const int *__restrict__ mptr,
                                                         const int *__restrict__ mem, const int *__restrict__ agg,
                                                         const int *__restrict__ rpc, int *__restrict__ colc,
                                                         int *__restrict__ dposc, TMc *__restrict__ valc,
                                                         double scale, int *__restrict__ err) {
  constexpr int NB = NV * NV, T = G::THREADS;
  __shared__ int ht[G::HASH];
  __shared__ int uj[G::COLS + 1], sj[G::COLS];
  __shared__ int ccode[G::CAND], cslot[G::CAND], corder[G::CAND];
  __shared__ int cnt, bad, nu;
  const int I = blockIdx.x, t = threadIdx.x;
  if (I >= nc) return;
  if (!gather_cols<G>(a, I, mptr, mem, agg, ht, &cnt, &bad)) {
    if (t == 0) atomicOr(err, 2);
    return;
  }
  // compact the table, then sort the columns by rank
  if (t == 0) nu = 0;
  __syncthreads();
  for (int h = t; h < G::HASH; h += T)
    if (ht[h] >= 0) uj[atomicAdd(&nu, 1)] = ht[h];
  __syncthreads();
  const int u = nu;
  for (int x = t; x < u; x += T) {
    const int v = uj[x];
    int r = 0;
    for (int y = 0; y < u; y++) r += uj[y] < v;
    sj[r] = v;
  }
  __syncthreads();
  const int base = rpc[I];
  for (int r = t; r < u; r += T) {
    colc[base + r] = sj[r];
    if (sj[r] == I) dposc[I] = r;
  }
  // candidates in member order: packed (row, entry) and the coarse slot of their column
  const int m0 = mptr[I], m1 = mptr[I + 1];
  int off = 0;
  for (int mi = m0; mi < m1; mi++) {
    const int i = mem[mi];
    const int L = a.len(i);
    for (int q = t; q < L; q += T) {
      const int j = a.col(i, q);
      const int J = j >= 0 ? agg[j] : -1;
      int slot = -1;
      if (J >= 0) {
        int lo = 0, hi = u - 1;
        while (lo <= hi) {
          const int mid = (lo + hi) >> 1;
          if (sj[mid] == J) { slot = mid; break; }
          if (sj[mid] < J) lo = mid + 1; else hi = mid - 1;
        }
      }
      ccode[off + q] = a.code(i, q);
      cslot[off + q] = slot;
    }
    off += L;
  }
  // slot starts (uj reused): counts, then an exclusive scan by one thread
  for (int r = t; r <= u; r += T) uj[r] = 0;
  __syncthreads();
  for (int cc = t; cc < off; cc += T)
    if (cslot[cc] >= 0) atomicAdd(&uj[cslot[cc]], 1);
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int r = 0; r <= u; r++) { const int v = uj[r]; uj[r] = acc; acc += v; }
  }
  __syncthreads();
  // stable placement: rank among earlier candidates of the same slot
  for (int cc = t; cc < off; cc += T) {
    const int sl = cslot[cc];
    if (sl < 0) continue;
    int k = 0;
    for (int y = 0; y < cc; y++) k += cslot[y] == sl;
    corder[uj[sl] + k] = ccode[cc];
  }
  __syncthreads();
  // block (slot r, entry e): its fine blocks in order
  for (int task = t; task < u * NB; task += T) {
    const int r = task / NB, e = task - r * NB;
    double acc = 0;
    for (int p = uj[r]; p < uj[r + 1]; p++) acc += a.vc(corder[p], e);
    valc[(size_t)(base + r) * NB + e] = (TMc)(acc * scale);
  }
}

// ---------------------------------------------------------------------------
// relaxation set-up
// ---------------------------------------------------------------------------
template <int NV, int KIND, typename Acc, typename TW>
__global__ __launch_bounds__(256) void k_relax_setup(Acc a, int n, double damping, TW *__restrict__ w,
                                                     int *__restrict__ err) {
  constexpr int NB = NV * NV;
  const int i = row_index();
  if (i >= n) return;
  const int dq = a.diag(i);
  if (KIND == kBlockJacobi) {
    double D[NB], Inv[NB];
    for (int e = 0; e < NB; e++) D[e] = a.v(i, dq, e);
    block_inverse<NV>(D, Inv);
    for (int e = 0; e < NB; e++) {
      const double x = damping * Inv[e];
      if (!isfinite(x)) atomicOr(err, 4);
      w[(size_t)i * NB + e] = (TW)x;
    }
  } else {
    for (int r = 0; r < NV; r++) {
      const double arr = a.v(i, dq, r * NV + r);
      double x;
      if (KIND == kJacobi) {
        x = damping * (1.0 / arr);
      } else {
        double den = 0;
        const int L = a.len(i);
        for (int q = 0; q < L; q++) {
          if (a.col(i, q) < 0) continue;
          for (int c = 0; c < NV; c++) {
            const double v = a.v(i, q, r * NV + c);
            den += v * v;
          }
        }
        x = (1.0 / den) * arr;
      }
      if (!isfinite(x)) atomicOr(err, 4);
      w[(size_t)i * NV + r] = (TW)x;
    }
  }
}

// ---------------------------------------------------------------------------
// cycle kernels (vectors [n][nv]; relaxation data in the matrix type TW)
// ---------------------------------------------------------------------------
// d = W_i t; ws (FP16 copies) divides by the row scale
template <int NV, int KIND, typename TW>
__device__ __forceinline__ void relax_apply(const TW *__restrict__ w, const float *__restrict__ ws, int i,
                                            const double *t, double *d) {
  if (KIND == kBlockJacobi) {
    const TW *W = w + (size_t)i * NV * NV;
    for (int r = 0; r < NV; r++) {
      double s = 0;
      for (int c = 0; c < NV; c++) s += to_d(W[r * NV + c]) * t[c];
      d[r] = s;
    }
  } else {
    for (int r = 0; r < NV; r++) d[r] = to_d(w[(size_t)i * NV + r]) * t[r];
  }
  if (ws) {
    const double inv = 1.0 / (double)ws[i];
    for (int r = 0; r < NV; r++) d[r] *= inv;
  }
}

// x = W f (one sweep from x = 0)
template <int NV, int KIND, typename TW, typename TV>
__global__ __launch_bounds__(256) void k_relax0(int n, const TW *__restrict__ w, const float *__restrict__ ws,
                                                const TV *__restrict__ f, TV *__restrict__ x) {
  const int i = row_index();
  if (i >= n) return;
  double t[NV], d[NV];
  for (int r = 0; r < NV; r++) t[r] = (double)f[(size_t)i * NV + r];
  relax_apply<NV, KIND>(w, ws, i, t, d);
  for (int r = 0; r < NV; r++) x[(size_t)i * NV + r] = (TV)d[r];
}

template <int NV, typename Acc, typename TV>
__device__ __forceinline__ void row_resid(const Acc &a, int i, const TV *__restrict__ f, const TV *__restrict__ x,
                                          double *t) {
  for (int r = 0; r < NV; r++) t[r] = (double)f[(size_t)i * NV + r];
  const int L = a.len(i);
  for (int q = 0; q < L; q++) {
    const int j = a.col(i, q);
    if (j < 0) continue;
    double xj[NV];
    for (int c = 0; c < NV; c++) xj[c] = (double)x[(size_t)j * NV + c];
    for (int r = 0; r < NV; r++) {
      double s = 0;
      for (int c = 0; c < NV; c++) s += a.v(i, q, r * NV + c) * xj[c];
      t[r] -= s;
    }
  }
}

// level 0 (ELL-BSR, entry-major): one thread per row, coalesced over rows
template <int NV, typename Acc, typename TV>
__global__ __launch_bounds__(256) void k_resid_lv(Acc a, int n, const TV *__restrict__ f, const TV *__restrict__ x,
                                                  TV *__restrict__ t) {
  const int i = row_index();
  if (i >= n) return;
  double r[NV];
  row_resid<NV>(a, i, f, x, r);
  for (int c = 0; c < NV; c++) t[(size_t)i * NV + c] = (TV)r[c];
}

// xout = xin + W (f - A xin)
template <int NV, int KIND, typename Acc, typename TW, typename TV>
__global__ __launch_bounds__(256) void k_relax(Acc a, int n, const TW *__restrict__ w, const float *__restrict__ ws,
                                               const TV *__restrict__ f, const TV *__restrict__ xin,
                                               TV *__restrict__ xout) {
  const int i = row_index();
  if (i >= n) return;
  double t[NV], d[NV];
  row_resid<NV>(a, i, f, xin, t);
  relax_apply<NV, KIND>(w, ws, i, t, d);
  for (int r = 0; r < NV; r++) xout[(size_t)i * NV + r] = (TV)((double)xin[(size_t)i * NV + r] + d[r]);
}

// row-scaled FP16 copies of the level-0 matrix and relaxation data: m_i is the
// largest |entry| of the diagonal block; Ah = A_i / m_i, Wh = W_i m_i
template <int NV, typename TM>
__global__ __launch_bounds__(256) void k_half_copy(const TM *__restrict__ A, const TM *__restrict__ W, int N, int K,
                                                   int wper, __half *__restrict__ Ah, __half *__restrict__ Wh,
                                                   float *__restrict__ rm, int *__restrict__ err) {
  constexpr int NB = NV * NV;
  const int i = row_index();
  if (i >= N) return;
  double m = 0;
  for (int e = 0; e < NB; e++) m = fmax(m, fabs(to_d(A[(size_t)e * N + i])));
  if (!(m > 0)) m = 1.0;
  const float mf = (float)m;
  rm[i] = mf;
  const double inv = 1.0 / (double)mf;
  bool ok = true;
  for (int q = 0; q <= K; q++)
    for (int e = 0; e < NB; e++) {
      const float v = (float)(to_d(A[((size_t)q * NB + e) * N + i]) * inv);
      ok &= fabsf(v) < 65000.0f;
      Ah[((size_t)q * NB + e) * N + i] = __float2half(v);
    }
  for (int e = 0; e < wper; e++) {
    const float v = (float)(to_d(W[(size_t)i * wper + e]) * (double)mf);
    ok &= fabsf(v) < 65000.0f;
    Wh[(size_t)i * wper + e] = __float2half(v);
  }
  if (!ok) atomicOr(err, 16);
}

// coarse levels (block CSR, blocks contiguous): 32 lanes per row, lane e of
// the group owns entry e of every block, so each block is one coalesced read.
// Row sums over the nv lanes of a block row by shuffles.  RELAX = 0 writes
// t = f - A x, RELAX = 1 writes x + W (f - A x).
// TS is the value type of the sweep copy (the matrix type, or FP16 with row
// scales rm: A_i = rm_i * val_i)
template <int NV, int KIND, int RELAX, typename TS, typename TM, typename TV>
__global__ __launch_bounds__(256) void k_csr_group(const int *__restrict__ rp, const int *__restrict__ cl,
                                                   const TS *__restrict__ val, const float *__restrict__ rm, int n,
                                                   const TM *__restrict__ w, const TV *__restrict__ f,
                                                   const TV *__restrict__ x, TV *__restrict__ out) {
  constexpr int NB = NV * NV;
  static_assert(NB <= 32, "one block per 32-lane group");
  const int i = (blockIdx.x * 256 + threadIdx.x) >> 5;
  const int l = threadIdx.x & 31, gbase = threadIdx.x & 32;   // lane in group, group's first lane in the wave
  const bool act = l < NB;
  const int r = act ? l / NV : 0, cc = act ? l - r * NV : 0;
  // blocks whose loads are in flight together: 8 or 16 hold more registers
  // and gave slower V-cycles (5 waves per SIMD instead of 8).  Neither a lane
  // per block row (12 rows a wave, 4 waves per SIMD) nor the row's entries
  // dealt over a whole wave was faster on the 300k-row level of the 2M wing
  constexpr int U = 4;
  // the row's own data, loaded first so that its latency overlaps the sweep's
  double fr = 0, xr = 0, wl = 0;
  if (i < n && act) {
    if (cc == 0) fr = (double)f[(size_t)i * NV + r];
    if (RELAX && cc == 0) xr = (double)x[(size_t)i * NV + r];
    if (RELAX && KIND == kBlockJacobi) wl = to_d(w[(size_t)i * NB + l]);
    if (RELAX && KIND != kBlockJacobi && cc == 0) wl = to_d(w[(size_t)i * NV + r]);
  }
  double acc = 0;
  if (i < n) {
    // the row's columns come in 32 at a time (one per lane) and reach the
    // group by shuffles; the block and x loads of U columns are issued before
    // any of them is used, and the products are summed in column order
    const int p0 = rp[i], len = rp[i + 1] - p0;
    for (int b0 = 0; b0 < len; b0 += 32) {
      const int nb = min(32, len - b0);
      const int Jl = l < nb ? cl[p0 + b0 + l] : 0;
      for (int q0 = 0; q0 < nb; q0 += U) {
        TS vq[U];
        TV xq[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int J = __shfl(Jl, gbase + ((q0 + u) & 31), 64);
          vq[u] = 0;
          xq[u] = 0;
          if (act && q0 + u < nb) {
            vq[u] = val[(size_t)(p0 + b0 + q0 + u) * NB + l];
            xq[u] = x[(size_t)J * NV + cc];
          }
        }
#pragma unroll
        for (int u = 0; u < U; u++)
          if (act && q0 + u < nb) acc += to_d(vq[u]) * (double)xq[u];
      }
    }
  }
  // every lane of the wave takes part in the shuffles
  double s = 0;
  for (int c = 0; c < NV; c++) s += __shfl(acc, gbase + r * NV + c, 64);
  if (rm && i < n) s *= (double)rm[i];
  double t = 0;
  if (i < n && act && cc == 0) t = fr - s;   // lane r*NV: row r of f - A x
  if (RELAX == 0) {
    if (i < n && act && cc == 0) out[(size_t)i * NV + r] = (TV)t;
    return;
  }
  double d;
  if (KIND == kBlockJacobi) {
    // lane r*NV + c: W[r][c] * t_c, t_c from lane c*NV, then the row sum
    const double tc = __shfl(t, gbase + cc * NV, 64);
    const double prod = (i < n && act) ? wl * tc : 0.0;
    d = 0;
    for (int c = 0; c < NV; c++) d += __shfl(prod, gbase + r * NV + c, 64);
  } else {
    d = (i < n && act && cc == 0) ? wl * t : 0.0;
  }
  if (i < n && act && cc == 0) out[(size_t)i * NV + r] = (TV)(xr + d);
}

// fc[I] = sum of t over the members of I: G lanes per coarse row (G a power
// of two <= 32, about the mean aggregate size), lane m takes members m, m + G,
// ...; the lanes' sums meet in a fixed butterfly
template <int NV, int G, typename TV>
__global__ __launch_bounds__(256) void k_restrict(int nc, const int *__restrict__ mptr, const int *__restrict__ mem,
                                                  const TV *__restrict__ t, TV *__restrict__ fc) {
  const int I = (blockIdx.x * 256 + threadIdx.x) / G, l = threadIdx.x & (G - 1);
  double s[NV];
  for (int r = 0; r < NV; r++) s[r] = 0;
  if (I < nc)
    for (int m = mptr[I] + l; m < mptr[I + 1]; m += G) {
      const int i = mem[m];
      for (int r = 0; r < NV; r++) s[r] += (double)t[(size_t)i * NV + r];
    }
  for (int o = G / 2; o > 0; o >>= 1)   // offsets below G stay inside the group
    for (int r = 0; r < NV; r++) s[r] += __shfl_xor(s[r], o, 64);
  if (I < nc && l == 0)
    for (int r = 0; r < NV; r++) fc[(size_t)I * NV + r] = (TV)s[r];
}

// row-scaled FP16 sweep copy of a coarse level: rm_i = largest |entry| of the
// diagonal block, hval = val / rm_i (err |= 64 when out of FP16 range)
template <int NV, typename TM>
__global__ __launch_bounds__(256) void k_half_csr(int n, const int *__restrict__ rp, const int *__restrict__ dpos,
                                                  const TM *__restrict__ val, __half *__restrict__ hval,
                                                  float *__restrict__ rm, int *__restrict__ err) {
  constexpr int NB = NV * NV;
  const int i = row_index();
  if (i >= n) return;
  double m = 0;
  for (int e = 0; e < NB; e++) m = fmax(m, fabs(to_d(val[(size_t)(rp[i] + dpos[i]) * NB + e])));
  if (!(m > 0)) m = 1.0;
  const float mf = (float)m;
  rm[i] = mf;
  const double inv = 1.0 / (double)mf;
  bool ok = true;
  for (size_t k = (size_t)rp[i] * NB; k < (size_t)rp[i + 1] * NB; k++) {
    const float v = (float)(to_d(val[k]) * inv);
    ok &= fabsf(v) < 65000.0f;
    hval[k] = __float2half(v);
  }
  if (!ok) atomicOr(err, 64);
}

template <int NV, typename TV>
__global__ __launch_bounds__(256) void k_prolong(int n, const int *__restrict__ agg, const TV *__restrict__ uc,
                                                 TV *__restrict__ x) {
  const int i = row_index();
  if (i >= n) return;
  const int g = agg[i];
  if (g < 0) return;
  for (int r = 0; r < NV; r++)
    x[(size_t)i * NV + r] = (TV)((double)x[(size_t)i * NV + r] + (double)uc[(size_t)g * NV + r]);
}

// u = inv f, one wave per row of the dense inverse (m x m, row-major)
template <typename TV>
__global__ __launch_bounds__(256) void k_dense_apply(int m, const double *__restrict__ inv, const TV *__restrict__ f,
                                                     TV *__restrict__ u) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= m) return;
  double s = 0;
  for (int c = lane; c < m; c += 64) s += inv[(size_t)row * m + c] * (double)f[c];
  s = wave_sum(s);
  if (lane == 0) u[row] = (TV)s;
}

// ---------------------------------------------------------------------------
// coarsest system: dense [A | I] (m x 2m) and Gauss-Jordan steps
// ---------------------------------------------------------------------------
template <int NV, typename Acc>
__global__ void k_densify(Acc a, int n, double *__restrict__ M) {
  const int m = n * NV;
  const int I = row_index();
  if (I >= n) return;
  for (int r = 0; r < NV; r++) {
    double *row = M + (size_t)(I * NV + r) * 2 * m;
    for (int c = 0; c < 2 * m; c++) row[c] = 0.0;
    row[m + I * NV + r] = 1.0;
  }
  const int L = a.len(I);
  for (int q = 0; q < L; q++) {
    const int J = a.col(I, q);
    if (J < 0) continue;
    for (int r = 0; r < NV; r++)
      for (int c = 0; c < NV; c++) M[(size_t)(I * NV + r) * 2 * m + J * NV + c] = a.v(I, q, r * NV + c);
  }
}

// one elimination step on column k: rows swapped with the pivot row (largest
// |entry| at or below k, first on ties), pivot row scaled, column k cleared.
// Reads Min, writes Mout (no in-place hazards across blocks).
__global__ __launch_bounds__(256) void k_gj_step(const double *__restrict__ Min, double *__restrict__ Mout, int m,
                                                 int k, int *__restrict__ err) {
  __shared__ double bv[256];
  __shared__ int bi[256];
  const int t = threadIdx.x;
  double best = -1.0;
  int bidx = k;
  for (int r = k + t; r < m; r += 256) {
    const double a = fabs(Min[(size_t)r * 2 * m + k]);
    if (a > best) { best = a; bidx = r; }
  }
  bv[t] = best;
  bi[t] = bidx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) {
      if (bv[t + o] > bv[t] || (bv[t + o] == bv[t] && bi[t + o] < bi[t])) { bv[t] = bv[t + o]; bi[t] = bi[t + o]; }
    }
    __syncthreads();
  }
  const int p = bi[0];
  const double piv = Min[(size_t)p * 2 * m + k];
  const long total = (long)m * 2 * m;
  for (long id = (long)blockIdx.x * 256 + t; id < total; id += (long)gridDim.x * 256) {
    const int r = (int)(id / (2 * m)), cidx = (int)(id - (long)r * 2 * m);
    if (piv == 0.0) {
      Mout[id] = Min[id];
      if (id == 0) atomicAdd(err, 1);
      continue;
    }
    const double pc = Min[(size_t)p * 2 * m + cidx] / piv;
    if (r == k) {
      Mout[id] = pc;
    } else {
      const int src = (r == p) ? k : r;
      Mout[id] = Min[(size_t)src * 2 * m + cidx] - Min[(size_t)src * 2 * m + k] * pc;
    }
  }
}

__global__ void k_take_inverse(int m, const double *__restrict__ M, double *__restrict__ inv) {
  const long total = (long)m * m;
  for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
    const int r = (int)(id / m), c = (int)(id - (long)r * m);
    inv[id] = M[(size_t)r * 2 * m + m + c];
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Device memory of a build comes from one arena sized by the previous build's
// total (first build: separate allocations), so the rebuild at every
// nonlinear step allocates nothing once the sizes have settled.
static int dmalloc(Amg *g, void **p, size_t bytes) {
  bytes = (std::max<size_t>(bytes, 16) + 255) & ~(size_t)255;
  g->arena_need += bytes;
  if (g->arena_off + bytes <= g->arena_cap) {
    *p = (char *)g->arena + g->arena_off;
    g->arena_off += bytes;
    return BES_OK;
  }
  BES_HIP(hipMalloc(p, bytes));
  g->owned.push_back(*p);
  return BES_OK;
}

static void free_levels(Amg *g) {
  for (void *p : g->owned) hipFree(p);
  g->owned.clear();
  g->arena_off = 0;
  g->lv.clear();
  g->scan_tmp.clear();
  g->dense = g->dense2 = g->inv = nullptr;
  g->Ah = g->Wh = nullptr;
  g->rm = nullptr;
  g->s = g->dia = nullptr;
  g->strong = g->sym = g->near = nullptr;
  g->m1 = nullptr;
  g->iflag = g->ibuf = g->ibuf2 = g->flags = nullptr;
  g->t0 = nullptr;
  g->nlev = 0;
  g->built_type = 0;
}

static inline int blocks_for(long n, int bs) { return (int)std::max<long>(1, (n + bs - 1) / bs); }

// exclusive scan of n ints (in -> out may alias); returns the total
static int scan_ex(bes_ctx *c, Amg *g, const int *in, int *out, int n, int depth, long *total) {
  const int nb = blocks_for(n, 1024);
  if ((int)g->scan_tmp.size() <= depth) {
    int *p = nullptr;
    // sized for level 0: nb <= ceil(N/1024) at depth 0, and shrinking by 1024 per depth
    const long cap = (long)blocks_for(g->lv[0].n + 1, 1024) + 16;
    int s = dmalloc(g, (void **)&p, (size_t)2 * cap * sizeof(int));
    if (s) return s;
    g->scan_tmp.push_back(p);
  }
  int *bsum = g->scan_tmp[depth];
  int *boff = bsum + blocks_for(g->lv[0].n + 1, 1024) + 16;
  hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(1024), 0, c->stream, in, out, n, bsum);
  BES_HIP(hipGetLastError());
  if (nb > 1) {
    int s = scan_ex(c, g, bsum, boff, nb, depth + 1, nullptr);
    if (s) return s;
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(1024), 0, c->stream, out, n, boff);
    BES_HIP(hipGetLastError());
    if (total) {
      int last[2];
      BES_HIP(hipMemcpyAsync(&last[0], boff + nb - 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      BES_HIP(hipMemcpyAsync(&last[1], bsum + nb - 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      BES_HIP(hipStreamSynchronize(c->stream));
      *total = (long)last[0] + last[1];
    }
  } else if (total) {
    int last;
    BES_HIP(hipMemcpyAsync(&last, bsum, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    BES_HIP(hipStreamSynchronize(c->stream));
    *total = last;
  }
  return BES_OK;
}

static int read_flags(bes_ctx *c, Amg *g, int *h) {
  BES_HIP(hipMemcpyAsync(h, g->flags, 3 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  return BES_OK;
}

// aggregation of level l (rows n, accessor a) into g->lv[l].agg; returns the count
template <int NV, typename Acc>
static int aggregate(bes_ctx *c, Amg *g, int l, const Acc &a, int *naggr) {
  const int bs = 256;
  Level &L = g->lv[l];
  const int n = L.n, nb = blocks_for(n, bs);
  const double e2 = (double)((float)g->prm.eps_strong * (float)g->prm.eps_strong);
  hipLaunchKernelGGL((k_pointwise<NV, Acc>), dim3(nb), dim3(bs), 0, c->stream, a, n, g->s, g->dia);
  hipLaunchKernelGGL((k_strong<Acc>), dim3(nb), dim3(bs), 0, c->stream, a, n, g->s, g->dia, e2, g->strong, L.state);
  BES_HIP(hipGetLastError());
  if (g->prm.sweep) {
    int *order = nullptr;
    if (l == 0) {   // the reference's cell order: step r takes device row iperm[r]
      std::vector<int> ip(c->iperm.begin(), c->iperm.end());
      int s = dmalloc(g, (void **)&order, (size_t)n * sizeof(int));
      if (s) return s;
      BES_HIP(hipMemcpyAsync(order, ip.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    }
    hipLaunchKernelGGL((k_row_sweep<Acc>), dim3(1), dim3(64), 0, c->stream, a, n, order, g->strong, L.agg, L.state,
                       g->ibuf2, g->flags + 3);
    BES_HIP(hipGetLastError());
    int h4[4];
    BES_HIP(hipMemcpyAsync(h4, g->flags, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    BES_HIP(hipStreamSynchronize(c->stream));
    *naggr = h4[3];
    return BES_OK;
  }
  hipLaunchKernelGGL((k_sym<Acc>), dim3(nb), dim3(bs), 0, c->stream, a, n, g->strong, L.state, g->sym);
  BES_HIP(hipGetLastError());
  const unsigned salt = 0x9e3779b9u * (unsigned)(l + 1);
  int rounds = 0;
  for (;; rounds++) {
    if (rounds >= kMaxRounds) { set_error("AMG aggregation did not finish"); return BES_ERR_SOLVE; }
    hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, c->stream, g->flags, 1L, 0);
    hipLaunchKernelGGL((k_mis_max1<Acc>), dim3(nb), dim3(bs), 0, c->stream, a, n, L.state, g->sym, salt, g->m1);
    hipLaunchKernelGGL((k_mis_max2<Acc>), dim3(nb), dim3(bs), 0, c->stream, a, n, L.state, g->sym, salt, g->m1);
    hipLaunchKernelGGL((k_mis_near<Acc>), dim3(nb), dim3(bs), 0, c->stream, a, n, L.state, g->sym, g->near);
    hipLaunchKernelGGL((k_mis_cover<Acc>), dim3(nb), dim3(bs), 0, c->stream, a, n, L.state, g->sym, g->near,
                       g->flags);
    BES_HIP(hipGetLastError());
    int h[3];
    int s = read_flags(c, g, h);
    if (s) return s;
    if (h[0] == 0) break;
  }
  hipLaunchKernelGGL(k_root_flag, dim3(nb), dim3(bs), 0, c->stream, n, L.state, g->iflag);
  long nr = 0;
  int s = scan_ex(c, g, g->iflag, g->ibuf, n, 0, &nr);
  if (s) return s;
  hipLaunchKernelGGL((k_assign1<Acc>), dim3(nb), dim3(bs), 0, c->stream, a, n, L.state, g->sym, g->ibuf, g->ibuf2);
  hipLaunchKernelGGL((k_assign2<Acc>), dim3(nb), dim3(bs), 0, c->stream, a, n, g->sym, g->s, g->ibuf2, L.agg,
                     g->flags + 1);
  BES_HIP(hipGetLastError());
  *naggr = (int)nr;
  return BES_OK;
}

// members of each aggregate, ascending
static int members(bes_ctx *c, Amg *g, int l, int nc) {
  Level &L = g->lv[l];
  const int n = L.n, bs = 256;
  int s;
  if ((s = dmalloc(g, (void **)&L.mptr, (size_t)(nc + 1) * sizeof(int))) ||
      (s = dmalloc(g, (void **)&L.mem, (size_t)n * sizeof(int))))
    return s;
  hipLaunchKernelGGL(k_fill_i, dim3(256), dim3(256), 0, c->stream, g->ibuf, (long)nc + 1, 0);
  hipLaunchKernelGGL(k_count, dim3(blocks_for(n, bs)), dim3(bs), 0, c->stream, n, L.agg, g->ibuf);
  long tot = 0;
  if ((s = scan_ex(c, g, g->ibuf, L.mptr, nc + 1, 0, &tot))) return s;
  hipLaunchKernelGGL(k_fill_i, dim3(256), dim3(256), 0, c->stream, g->ibuf, (long)nc + 1, 0);
  hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n, bs)), dim3(bs), 0, c->stream, n, L.agg, L.mptr, g->ibuf, L.mem);
  hipLaunchKernelGGL(k_sort_members, dim3(blocks_for(nc, bs)), dim3(bs), 0, c->stream, nc, L.mptr, L.mem);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

template <int NV, typename TM, typename Acc>
static int galerkin(bes_ctx *c, Amg *g, int l, const Acc &a, int nc) {
  constexpr int NB = NV * NV;
  Level &L = g->lv[l], &C = g->lv[l + 1];
  int s, h[3];
  // column counts with the small configuration; a level with a row that does
  // not fit goes to the medium one, then to the large one
  int cfg = 0;
  hipLaunchKernelGGL((k_gal_count<GalSmall, Acc>), dim3(nc), dim3(GalSmall::THREADS), 0, c->stream, a, nc, L.mptr,
                     L.mem, L.agg, g->ibuf, g->flags + 1, 8);
  BES_HIP(hipGetLastError());
  if ((s = read_flags(c, g, h))) return s;
  if (h[1] & 8) {
    cfg = 1;
    hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, c->stream, g->flags + 1, 1L, h[1] & ~8);
    hipLaunchKernelGGL((k_gal_count<GalMedium, Acc>), dim3(nc), dim3(GalMedium::THREADS), 0, c->stream, a, nc,
                       L.mptr, L.mem, L.agg, g->ibuf, g->flags + 1, 8);
    BES_HIP(hipGetLastError());
    if ((s = read_flags(c, g, h))) return s;
    if (h[1] & 8) {
      cfg = 2;
      hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, c->stream, g->flags + 1, 1L, h[1] & ~8);
      hipLaunchKernelGGL((k_gal_count<GalLarge, Acc>), dim3(nc), dim3(GalLarge::THREADS), 0, c->stream, a, nc,
                         L.mptr, L.mem, L.agg, g->ibuf, g->flags + 1, 2);
      BES_HIP(hipGetLastError());
    }
  }
  if ((s = dmalloc(g, (void **)&C.rp, (size_t)(nc + 1) * sizeof(int)))) return s;
  hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, c->stream, g->ibuf + nc, 1L, 0);
  long nnzb = 0;
  if ((s = scan_ex(c, g, g->ibuf, C.rp, nc + 1, 0, &nnzb))) return s;
  if ((s = read_flags(c, g, h))) return s;
  if (h[1] & 2) { set_error("AMG coarse row exceeds the gather limits"); return BES_ERR_SOLVE; }
  C.n = nc;
  C.nnzb = nnzb;
  if ((s = dmalloc(g, (void **)&C.col, (size_t)nnzb * sizeof(int))) ||
      (s = dmalloc(g, (void **)&C.dpos, (size_t)nc * sizeof(int))) ||
      (s = dmalloc(g, &C.val, (size_t)nnzb * NB * sizeof(TM))))
    return s;
  hipLaunchKernelGGL(k_fill_i, dim3(256), dim3(256), 0, c->stream, C.dpos, (long)nc, 0);
  const double scale = (double)(1.0f / (float)g->prm.over_interp);
  if (cfg == 2)
    hipLaunchKernelGGL((k_gal_fill<GalLarge, NV, Acc, TM>), dim3(nc), dim3(GalLarge::THREADS), 0, c->stream, a, nc,
                       L.mptr, L.mem, L.agg, C.rp, C.col, C.dpos, (TM *)C.val, scale, g->flags + 1);
  else if (cfg == 1)
    hipLaunchKernelGGL((k_gal_fill<GalMedium, NV, Acc, TM>), dim3(nc), dim3(GalMedium::THREADS), 0, c->stream, a,
                       nc, L.mptr, L.mem, L.agg, C.rp, C.col, C.dpos, (TM *)C.val, scale, g->flags + 1);
  else
    hipLaunchKernelGGL((k_gal_fill<GalSmall, NV, Acc, TM>), dim3(nc), dim3(GalSmall::THREADS), 0, c->stream, a, nc,
                       L.mptr, L.mem, L.agg, C.rp, C.col, C.dpos, (TM *)C.val, scale, g->flags + 1);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

template <int NV, typename TM, typename Acc>
static int relax_setup(bes_ctx *c, Amg *g, int l, const Acc &a) {
  Level &L = g->lv[l];
  const int n = L.n, bs = 256, nb = blocks_for(n, bs);
  const int per = g->prm.relax == kBlockJacobi ? NV * NV : NV;
  int s = dmalloc(g, &L.w, (size_t)n * per * sizeof(TM));
  if (s) return s;
  TM *w = (TM *)L.w;
  if (g->prm.relax == kBlockJacobi)
    hipLaunchKernelGGL((k_relax_setup<NV, kBlockJacobi, Acc, TM>), dim3(nb), dim3(bs), 0, c->stream, a, n,
                       g->prm.damping, w, g->flags + 1);
  else if (g->prm.relax == kSpai0)
    hipLaunchKernelGGL((k_relax_setup<NV, kSpai0, Acc, TM>), dim3(nb), dim3(bs), 0, c->stream, a, n, g->prm.damping,
                       w, g->flags + 1);
  else
    hipLaunchKernelGGL((k_relax_setup<NV, kJacobi, Acc, TM>), dim3(nb), dim3(bs), 0, c->stream, a, n, g->prm.damping,
                       w, g->flags + 1);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

template <int NV, typename TM>
static EllAcc<TM, NV> ell(bes_ctx *c, Impl *im) {
  return EllAcc<TM, NV>{(const TM *)im->A, c->dm.nbr, im->fd.mrev, c->N, c->k};
}
template <int NV, typename TM>
static CsrAcc<TM, NV> csr(const Level &L) {
  return CsrAcc<TM, NV>{L.rp, L.col, L.dpos, (const TM *)L.val};
}

template <int NV, typename TM>
static int build_t(bes_ctx *c, Impl *im, Amg *g) {
  const int N = c->N, K = c->k;
  const size_t es = im->prec == 32 ? sizeof(float) : sizeof(double);
  if (N >= (1 << 28)) { set_error("AMG: at most 2^28 cells"); return BES_ERR_ARG; }
  free_levels(g);
  if (g->arena_need > g->arena_cap) {
    if (g->arena) BES_HIP(hipFree(g->arena));
    g->arena = nullptr;
    g->arena_cap = 0;
    const size_t cap = g->arena_need + g->arena_need / 8;
    BES_HIP(hipMalloc(&g->arena, cap));
    g->arena_cap = cap;
  }
  g->arena_need = 0;
  int s;
  const size_t ent0 = (size_t)(K + 1) * N;
  g->lv.resize(1);
  g->lv[0].n = N;
  g->lv[0].nnzb = N + 2 * im->ninternal;
  if ((s = dmalloc(g, (void **)&g->s, ent0 * sizeof(double))) || (s = dmalloc(g, (void **)&g->dia, N * sizeof(double))) ||
      (s = dmalloc(g, (void **)&g->strong, ent0)) || (s = dmalloc(g, (void **)&g->sym, ent0)) ||
      (s = dmalloc(g, (void **)&g->near, N)) || (s = dmalloc(g, (void **)&g->m1, N * sizeof(unsigned long long))) ||
      (s = dmalloc(g, (void **)&g->iflag, (N + 1) * sizeof(int))) ||
      (s = dmalloc(g, (void **)&g->ibuf, (N + 1) * sizeof(int))) ||
      (s = dmalloc(g, (void **)&g->ibuf2, (N + 1) * sizeof(int))) ||
      (s = dmalloc(g, (void **)&g->flags, 4 * sizeof(int))) ||
      (s = dmalloc(g, &g->t0, (size_t)N * NV * es)))
    return s;
  hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, c->stream, g->flags, 4L, 0);
  g->direct = false;
  for (int l = 0;; l++) {
    Level &L = g->lv[l];
    const int n = L.n;
    if (l > 0 &&
        ((s = dmalloc(g, &L.f, (size_t)n * NV * es)) || (s = dmalloc(g, &L.u, (size_t)n * NV * es)) ||
         (s = dmalloc(g, &L.t, (size_t)n * NV * es)) || (s = dmalloc(g, &L.xa, (size_t)n * NV * es))))
      return s;
    if ((s = dmalloc(g, (void **)&L.agg, n * sizeof(int))) || (s = dmalloc(g, (void **)&L.state, n)))
      return s;
    s = l == 0 ? relax_setup<NV, TM>(c, g, l, ell<NV, TM>(c, im)) : relax_setup<NV, TM>(c, g, l, csr<NV, TM>(L));
    if (s) return s;
    if (l == 0 && im->prec == 32) {
      const int wper = g->prm.relax == kBlockJacobi ? NV * NV : NV;
      if ((s = dmalloc(g, (void **)&g->Ah, (size_t)(K + 1) * NV * NV * N * sizeof(__half))) ||
          (s = dmalloc(g, (void **)&g->Wh, (size_t)N * wper * sizeof(__half))) ||
          (s = dmalloc(g, (void **)&g->rm, (size_t)N * sizeof(float))))
        return s;
      hipLaunchKernelGGL((k_half_copy<NV, TM>), dim3(blocks_for(N, 256)), dim3(256), 0, c->stream,
                         (const TM *)im->A, (const TM *)L.w, N, K, wper, g->Ah, g->Wh, g->rm, g->flags + 1);
      BES_HIP(hipGetLastError());
    }
    const long rows = (long)n * NV;
    if (rows <= g->prm.coarse_enough) {   // coarsest: dense inverse
      g->direct = rows <= kDenseMax;
      break;
    }
    if (l + 1 >= kMaxLevels) break;
    int nc = 0;
    s = l == 0 ? aggregate<NV>(c, g, l, ell<NV, TM>(c, im), &nc) : aggregate<NV>(c, g, l, csr<NV, TM>(L), &nc);
    if (s) return s;
    int h[3];
    if ((s = read_flags(c, g, h))) return s;
    if (h[1] & 1) { set_error("AMG aggregation left a row unassigned"); return BES_ERR_SOLVE; }
    if (nc == 0 || nc >= n) break;   // nothing to coarsen: this level is relaxed as the coarsest
    if ((s = members(c, g, l, nc))) return s;
    g->lv.emplace_back();
    Level &Lr = g->lv[l];   // emplace_back may move the vector
    s = l == 0 ? galerkin<NV, TM>(c, g, l, ell<NV, TM>(c, im), nc) : galerkin<NV, TM>(c, g, l, csr<NV, TM>(Lr), nc);
    if (s) return s;
  }
  g->nlev = (int)g->lv.size();
  if (im->prec == 32)
    for (int l = 1; l < g->nlev; l++) {
      if (l == g->nlev - 1 && g->direct) break;   // the coarsest is applied as a dense inverse
      Level &Lh = g->lv[l];
      if ((s = dmalloc(g, (void **)&Lh.hval, (size_t)Lh.nnzb * NV * NV * sizeof(__half))) ||
          (s = dmalloc(g, (void **)&Lh.hrm, (size_t)Lh.n * sizeof(float))))
        return s;
      hipLaunchKernelGGL((k_half_csr<NV, TM>), dim3(blocks_for(Lh.n, 256)), dim3(256), 0, c->stream, Lh.n, Lh.rp,
                         Lh.dpos, (const TM *)Lh.val, Lh.hval, Lh.hrm, g->flags + 1);
      int hf[3];
      if ((s = read_flags(c, g, hf))) return s;
      if (hf[1] & 64) {   // out of FP16 range: this level sweeps its full-precision values
        Lh.hval = nullptr;
        Lh.hrm = nullptr;
        hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, c->stream, g->flags + 1, 1L, hf[1] & ~64);
      }
    }
  if (g->direct) {
    Level &Cl = g->lv.back();
    const int m = Cl.n * NV;
    g->ndense = m;
    if ((s = dmalloc(g, (void **)&g->dense, (size_t)m * 2 * m * sizeof(double))) ||
        (s = dmalloc(g, (void **)&g->dense2, (size_t)m * 2 * m * sizeof(double))) ||
        (s = dmalloc(g, (void **)&g->inv, (size_t)m * m * sizeof(double))))
      return s;
    if (g->nlev == 1)
      hipLaunchKernelGGL((k_densify<NV, EllAcc<TM, NV>>), dim3(blocks_for(Cl.n, 64)), dim3(64), 0, c->stream,
                         ell<NV, TM>(c, im), Cl.n, g->dense);
    else
      hipLaunchKernelGGL((k_densify<NV, CsrAcc<TM, NV>>), dim3(blocks_for(Cl.n, 64)), dim3(64), 0, c->stream,
                         csr<NV, TM>(Cl), Cl.n, g->dense);
    double *a = g->dense, *b = g->dense2;
    const int gb = blocks_for((long)m * 2 * m, 256);
    for (int k = 0; k < m; k++) {
      hipLaunchKernelGGL(k_gj_step, dim3(gb), dim3(256), 0, c->stream, a, b, m, k, g->flags + 2);
      std::swap(a, b);
    }
    hipLaunchKernelGGL(k_take_inverse, dim3(blocks_for((long)m * m, 256)), dim3(256), 0, c->stream, m, a, g->inv);
    BES_HIP(hipGetLastError());
  }
  int h[3];
  if ((s = read_flags(c, g, h))) return s;
  if (h[1] & 4) { set_error("AMG relaxation has a singular diagonal"); return BES_ERR_SOLVE; }
  if (h[1] & 16) g->Ah = g->Wh = nullptr, g->rm = nullptr;   // out of FP16 range: full-precision level 0
  if (h[2]) { set_error("AMG coarsest system is singular"); return BES_ERR_SOLVE; }
  g->built_type = (int)sizeof(TM);
  return BES_OK;
}

// ---------------------------------------------------------------------------
// V-cycle
// ---------------------------------------------------------------------------
template <int NV, typename TW, typename TV>
static void launch_relax0_t(bes_ctx *c, int kind, int n, const TW *w, const float *ws, const TV *f, TV *x) {
  const int nb = blocks_for(n, 256);
  if (kind == kBlockJacobi)
    hipLaunchKernelGGL((k_relax0<NV, kBlockJacobi, TW, TV>), dim3(nb), dim3(256), 0, c->stream, n, w, ws, f, x);
  else
    hipLaunchKernelGGL((k_relax0<NV, kJacobi, TW, TV>), dim3(nb), dim3(256), 0, c->stream, n, w, ws, f, x);
}

// x = W f on level l (level 0 of an FP32 solve from the FP16 copies)
template <int NV, typename TM, typename TV>
static void launch_relax0(bes_ctx *c, Amg *g, int l, const TV *f, TV *x) {
  const Level &L = g->lv[l];
  if (l == 0 && g->Ah) launch_relax0_t<NV, __half, TV>(c, g->prm.relax, L.n, g->Wh, g->rm, f, x);
  else launch_relax0_t<NV, TM, TV>(c, g->prm.relax, L.n, (const TM *)L.w, nullptr, f, x);
}

template <int NV, typename Acc, typename TW, typename TV>
static void launch_ell_sweep(bes_ctx *c, int kind, const Acc &a, int n, bool relax, const TW *w, const float *ws,
                             const TV *f, const TV *x, TV *out) {
  const int nb = blocks_for(n, 256);
  if (!relax)
    hipLaunchKernelGGL((k_resid_lv<NV, Acc, TV>), dim3(nb), dim3(256), 0, c->stream, a, n, f, x, out);
  else if (kind == kBlockJacobi)
    hipLaunchKernelGGL((k_relax<NV, kBlockJacobi, Acc, TW, TV>), dim3(nb), dim3(256), 0, c->stream, a, n, w, ws, f,
                       x, out);
  else
    hipLaunchKernelGGL((k_relax<NV, kJacobi, Acc, TW, TV>), dim3(nb), dim3(256), 0, c->stream, a, n, w, ws, f, x,
                       out);
}

template <int NV, typename TS, typename TM, typename TV>
static void launch_csr_sweep(bes_ctx *c, int kind, const Level &L, bool relax, const TS *val, const float *rm,
                             const TM *w, const TV *f, const TV *x, TV *out) {
  const int n = L.n, nb = blocks_for((long)n * 32, 256);
  if (!relax)
    hipLaunchKernelGGL((k_csr_group<NV, kJacobi, 0, TS, TM, TV>), dim3(nb), dim3(256), 0, c->stream, L.rp, L.col, val,
                       rm, n, w, f, x, out);
  else if (kind == kBlockJacobi)
    hipLaunchKernelGGL((k_csr_group<NV, kBlockJacobi, 1, TS, TM, TV>), dim3(nb), dim3(256), 0, c->stream, L.rp,
                       L.col, val, rm, n, w, f, x, out);
  else
    hipLaunchKernelGGL((k_csr_group<NV, kJacobi, 1, TS, TM, TV>), dim3(nb), dim3(256), 0, c->stream, L.rp, L.col, val,
                       rm, n, w, f, x, out);
}

// t = f - A x (relax = false) or xout = x + W (f - A x) (relax = true) on level l
template <int NV, typename TM, typename TV>
static void launch_sweep(bes_ctx *c, Impl *im, Amg *g, int l, bool relax, const TV *f, const TV *x, TV *out) {
  const Level &L = g->lv[l];
  const int n = L.n, kind = g->prm.relax;
  const TM *w = (const TM *)L.w;
  if (l == 0) {
    if (g->Ah)
      launch_ell_sweep<NV>(c, kind, EllAccH<NV>{g->Ah, g->rm, c->dm.nbr, c->N, c->k}, n, relax,
                           (const __half *)g->Wh, (const float *)g->rm, f, x, out);
    else
      launch_ell_sweep<NV>(c, kind, ell<NV, TM>(c, im), n, relax, w, (const float *)nullptr, f, x, out);
    return;
  }
  if (L.hval) launch_csr_sweep<NV>(c, kind, L, relax, (const __half *)L.hval, L.hrm, w, f, x, out);
  else launch_csr_sweep<NV>(c, kind, L, relax, (const TM *)L.val, (const float *)nullptr, w, f, x, out);
}

template <int NV, typename TM, typename TV>
static void cycle(bes_ctx *c, Impl *im, Amg *g, int l, const TV *f, TV *x, TV *xa, TV *t) {
  const Level &L = g->lv[l];
  const int n = L.n;   // SPAI0 and damped Jacobi share the scalar kernels
  const int nb = blocks_for(n, 256);
  if (l == g->nlev - 1) {
    if (g->direct) {
      hipLaunchKernelGGL((k_dense_apply<TV>), dim3(blocks_for(g->ndense, 4)), dim3(256), 0, c->stream, g->ndense,
                         g->inv, f, x);
    } else {
      launch_relax0<NV, TM, TV>(c, g, l, f, xa);
      launch_sweep<NV, TM, TV>(c, im, g, l, true, f, xa, x);
    }
    return;
  }
  const Level &C = g->lv[l + 1];
  launch_relax0<NV, TM, TV>(c, g, l, f, xa);
  launch_sweep<NV, TM, TV>(c, im, g, l, false, f, xa, t);
  if (n <= 12L * C.n)   // small aggregates: 8 lanes each
    hipLaunchKernelGGL((k_restrict<NV, 8, TV>), dim3(blocks_for((long)C.n * 8, 256)), dim3(256), 0, c->stream, C.n,
                       L.mptr, L.mem, t, (TV *)C.f);
  else
    hipLaunchKernelGGL((k_restrict<NV, 32, TV>), dim3(blocks_for((long)C.n * 32, 256)), dim3(256), 0, c->stream,
                       C.n, L.mptr, L.mem, t, (TV *)C.f);
  cycle<NV, TM, TV>(c, im, g, l + 1, (const TV *)C.f, (TV *)C.u, (TV *)C.xa, (TV *)C.t);
  hipLaunchKernelGGL((k_prolong<NV, TV>), dim3(nb), dim3(256), 0, c->stream, n, L.agg, (const TV *)C.u, xa);
  launch_sweep<NV, TM, TV>(c, im, g, l, true, f, xa, x);
}

static double level_bytes(const Amg *g, int nv, size_t es_m, size_t es_v) {
  // every level: its matrix twice (residual and post-sweep), relaxation data
  // twice, about eight vector passes; the dense inverse once
  double b = 0;
  for (int l = 0; l < g->nlev; l++) {
    const Level &L = g->lv[l];
    const double per = g->prm.relax == kBlockJacobi ? nv * nv : nv;
    const double em = ((l == 0 && g->Ah) || (l > 0 && L.hval)) ? 2.0 : (double)es_m;
    b += 2.0 * L.nnzb * nv * nv * em + 2.0 * L.n * per * em + 8.0 * L.n * nv * es_v;
  }
  if (g->direct) b += (double)g->ndense * g->ndense * 8.0;
  return b;
}

}  // namespace amg

using namespace amg;

int amg_build(bes_ctx *c, Impl *im) {
  if (!im->amg) im->amg = new Amg();
  Amg *g = amg_of(im);
  // a lagged hierarchy: reused for rebuild - 1 solves after the one that built
  // it (the cycle then preconditions a nearby matrix; GMRES still applies the
  // current one)
  const int type = im->prec == 32 ? 4 : 8;
  if (g->built_type == type && !g->force && g->reused + 1 < g->prm.rebuild) {
    g->reused++;
    return BES_OK;
  }
  g->reused = 0;
  g->force = false;
  g->builds++;
  const double bytes = (double)c->N * (c->k + 1) * c->nv * c->nv * (im->prec == 32 ? 4.0 : 8.0) * 3.0;
  ProfScope ps(c, kProfFactor, bytes);
  if (im->prec == 32) return c->nd == 2 ? build_t<4, float>(c, im, g) : build_t<5, float>(c, im, g);
  return c->nd == 2 ? build_t<4, double>(c, im, g) : build_t<5, double>(c, im, g);
}

template <typename TV>
int amg_apply(bes_ctx *c, Impl *im, const TV *r, TV *z, TV *tmp) {
  Amg *g = amg_of(im);
  if (!g || g->built_type != (im->prec == 32 ? 4 : 8)) { set_error("AMG hierarchy not built"); return BES_ERR_ARG; }
  ProfScope ps(c, kProfPrecond, level_bytes(g, c->nv, im->prec == 32 ? 4 : 8, sizeof(TV)));
  if (im->prec == 32) {
    if (c->nd == 2) cycle<4, float, TV>(c, im, g, 0, r, z, tmp, (TV *)g->t0);
    else cycle<5, float, TV>(c, im, g, 0, r, z, tmp, (TV *)g->t0);
  } else {
    if (c->nd == 2) cycle<4, double, TV>(c, im, g, 0, r, z, tmp, (TV *)g->t0);
    else cycle<5, double, TV>(c, im, g, 0, r, z, tmp, (TV *)g->t0);
  }
  BES_HIP(hipGetLastError());
  return BES_OK;
}
template int amg_apply<double>(bes_ctx *, Impl *, const double *, double *, double *);
template int amg_apply<float>(bes_ctx *, Impl *, const float *, float *, float *);

void amg_release(Impl *im) {
  if (!im->amg) return;
  free_levels(amg_of(im));
  if (amg_of(im)->arena) hipFree(amg_of(im)->arena);
  delete amg_of(im);
  im->amg = nullptr;
}

}  // namespace bes

using namespace bes;

extern "C" {

int bes_amg_set_params(bes_ctx *c, double eps_strong, double over_interp, double damping, int coarse_enough,
                       int relax, int rebuild_interval) {
  Impl *im = c ? (Impl *)c->impl : nullptr;
  if (!im || !(eps_strong > 0 && eps_strong < 1) || !(over_interp >= 1) || !(damping > 0) || coarse_enough < 1 ||
      relax < 0 || relax > 2 || rebuild_interval < 1) {
    set_error("call bes_implicit_setup first; 0 < eps_strong < 1, over_interp >= 1, damping > 0, relax 0-2, "
              "rebuild_interval >= 1");
    return BES_ERR_ARG;
  }
  if (!im->amg) im->amg = new Amg();
  Params &p = amg_of(im)->prm;
  p.eps_strong = eps_strong;
  p.over_interp = over_interp;
  p.damping = damping;
  p.coarse_enough = coarse_enough;
  p.relax = relax;
  p.rebuild = rebuild_interval;
  amg_of(im)->force = true;
  im->factored = false;
  return BES_OK;
}

int bes_amg_set_aggregation(bes_ctx *c, int mode) {
  Impl *im = c ? (Impl *)c->impl : nullptr;
  if (!im || mode < 0 || mode > 1) { set_error("call bes_implicit_setup first / mode 0 or 1"); return BES_ERR_ARG; }
  if (!im->amg) im->amg = new Amg();
  amg_of(im)->prm.sweep = mode;
  amg_of(im)->force = true;
  im->factored = false;
  return BES_OK;
}

int bes_amg_info(bes_ctx *c, int *nlevels, long *rows, long *nnzb, int *direct, long *builds, int maxlev) {
  Impl *im = c ? (Impl *)c->impl : nullptr;
  Amg *g = im ? amg_of(im) : nullptr;
  if (!g || !g->built_type) { set_error("AMG hierarchy not built"); return BES_ERR_ARG; }
  if (nlevels) *nlevels = g->nlev;
  if (builds) *builds = g->builds;
  for (int l = 0; l < g->nlev && l < maxlev; l++) {
    if (rows) rows[l] = g->lv[l].n;
    if (nnzb) nnzb[l] = g->lv[l].nnzb;
  }
  if (direct) *direct = g->direct ? 1 : 0;
  return BES_OK;
}

int bes_amg_level(bes_ctx *c, int level, long *agg, unsigned char *state, long *rowptr, long *col, double *val) {
  Impl *im = c ? (Impl *)c->impl : nullptr;
  Amg *g = im ? amg_of(im) : nullptr;
  if (!g || !g->built_type || level < 0 || level >= g->nlev) { set_error("AMG hierarchy not built / level"); return BES_ERR_ARG; }
  BES_HIP(hipStreamSynchronize(c->stream));
  const Level &L = g->lv[level];
  const int n = L.n, NB = c->nv * c->nv;
  if (agg || state) {
    const bool has = level < g->nlev - 1;
    std::vector<int> a(n, -1);
    std::vector<unsigned char> st(n, 0);
    if (has) {
      BES_HIP(hipMemcpy(a.data(), L.agg, n * sizeof(int), hipMemcpyDeviceToHost));
      BES_HIP(hipMemcpy(st.data(), L.state, n, hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < n; i++) {
      const int row = level == 0 ? (int)c->perm[i] : i;   // level 0 in the reference's cell order
      if (agg) agg[row] = has ? a[i] : -1;
      if (state) state[row] = st[i];
    }
  }
  if (level > 0 && (rowptr || col || val)) {
    std::vector<int> rp(n + 1), cl(L.nnzb);
    BES_HIP(hipMemcpy(rp.data(), L.rp, (n + 1) * sizeof(int), hipMemcpyDeviceToHost));
    BES_HIP(hipMemcpy(cl.data(), L.col, L.nnzb * sizeof(int), hipMemcpyDeviceToHost));
    if (rowptr) for (int i = 0; i <= n; i++) rowptr[i] = rp[i];
    if (col) for (long p = 0; p < L.nnzb; p++) col[p] = cl[p];
    if (val) {
      const size_t cnt = (size_t)L.nnzb * NB;
      if (g->built_type == 4) {
        std::vector<float> v(cnt);
        BES_HIP(hipMemcpy(v.data(), L.val, cnt * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < cnt; q++) val[q] = v[q];
      } else {
        BES_HIP(hipMemcpy(val, L.val, cnt * sizeof(double), hipMemcpyDeviceToHost));
      }
    }
  }
  return BES_OK;
}

}  // extern "C"
