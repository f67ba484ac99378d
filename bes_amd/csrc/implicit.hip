// Implicit path of the BES solver core on MI355X (gfx950).
//
// One backward-Euler pseudo-time step (SolverImplicit.cpp:293-345):
//   k_timestep + k_gradlim          (explicit.hip)
//   k_face_jacobian   first-order flux Jacobians dF/dU_L, dF/dU_R per face by
//                     one-direction dual numbers with the reference's Eigen AD
//                     rules (FluxJacobianWrapper.tpp:32-82), one thread per
//                     (face, input column)
//   k_assemble        row-owned assembly: second-order fluxes give the right-
//                     hand side b = -dt/V sum +-F*A; the block row gets
//                     I + sum nV x (dt/V*A*J) in the reference's accumulation
//                     order (SolverImplicit.cpp:358-507, CoefficientMatrix.cpp:
//                     166-204, the nV-fold repeat included).  No atomics.
//   GMRES(m)          right-preconditioned, classical Gram-Schmidt with DGKS
//                     re-orthogonalisation, deterministic two-pass reductions;
//                     replaces the reference's AMGCL solve (CoefficientMatrix.cpp:
//                     210-325) with relative tolerance semantics (||r|| <= tol*||b||)
//   preconditioners   block Jacobi, or block ILU(0) in a multicolour order
//                     (colours of the cell adjacency graph: every colour is
//                     one parallel sweep of the triangular solves)
//   k_update          U += x, the implicit update (SolverImplicit.cpp:509-528)
// Matrix layout ("ELL-BSR"): slot 0 = diagonal block, slot 1+m = the block of
// the neighbour through local face m; entries stored [slot][e][N] so a wave
// reads 64 consecutive rows of one entry in one coalesced load.  Matrix and
// Krylov vectors are FP64, or FP32 for the mixed-precision inner solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bes_hip.h"
#include "context.hpp"
#include "flux.hpp"
#include "implicit.hpp"
#include "kernels_common.hpp"

namespace bes {

int launch_timestep(bes_ctx *c, const double *U, double cfl, int local);
int launch_face_flux(bes_ctx *c);
int launch_gradlim(bes_ctx *c, const double *U, double *Gout, double *Pout);
int state_to_host(bes_ctx *c, const double *dsoa, int ncomp, double *host);
int state_to_device(bes_ctx *c, const double *host, int ncomp, double *dsoa);
namespace lp {
int lp_build(bes_ctx *c, Impl *im);
int lp_release(Impl *im);
}  // namespace lp

// ---------------------------------------------------------------------------
// Jacobians
// ---------------------------------------------------------------------------
struct FinfI {
  double p[6];
};

template <int ND>
__global__ __launch_bounds__(256) void k_face_jacobian(const double *__restrict__ U, int N, FaceDev fd,
                                                       GasConst gc, FinfI fi, double *__restrict__ JF) {
  constexpr int NV = ND + 2, NB = NV * NV;
  // groups of 64 faces x 2nv input directions: the lanes of a wave take 64
  // consecutive faces of one direction (coalesced entry-major stores), and the
  // group's waves reuse the same face states through the cache
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int grp = t / (64 * 2 * NV), rem = t - grp * (64 * 2 * NV);
  const int dir = rem >> 6, f = grp * 64 + (rem & 63);
  if (f >= fd.F) return;
  const int o = fd.fo[f], j = fd.fn[f];
  if (j < 0 && dir >= NV) return;   // boundary faces: left Jacobian only
  double nn[ND];
#pragma unroll
  for (int d = 0; d < ND; d++) nn[d] = fd.fnrm[(size_t)d * fd.F + f];
  Dual ul[NV], ur[NV + 1], F[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) ul[v] = Dual(U[(size_t)v * N + o], dir == v ? 1.0 : 0.0);
  if (j >= 0) {
#pragma unroll
    for (int v = 0; v < NV; v++) ur[v] = Dual(U[(size_t)v * N + j], dir == NV + v ? 1.0 : 0.0);
    roe_flux<ND, Dual>(gc, ul, ur, nn, F);
  } else if (j == -1) {
    slip_flux<ND, Dual>(gc, ul, nn, F);
  } else {
#pragma unroll
    for (int v = 0; v < NV + 1; v++) ur[v] = Dual(fi.p[v]);
    freestream_flux<ND, Dual>(gc, ul, ur, nn, F);
  }
  const int side = dir < NV ? 0 : 1, col = dir < NV ? dir : dir - NV;
#pragma unroll
  for (int r = 0; r < NV; r++) JF[((size_t)side * NB + r * NV + col) * fd.F + f] = F[r].d;
}

// ---------------------------------------------------------------------------
// assembly: right-hand side from second-order fluxes, blocks from JF
// ---------------------------------------------------------------------------
template <int ND, typename TM>
__global__ __launch_bounds__(256) void k_assemble(
    const double *__restrict__ FF, const double *__restrict__ dt, const double *__restrict__ dtg, DevMesh dm,
    FaceDev fd, const double *__restrict__ JF, int repeat, double *__restrict__ b, TM *__restrict__ A,
    unsigned long long *__restrict__ linf_b, unsigned long long *__restrict__ linf_r) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int N = dm.N;
  const int i = cell_index();
  double bi[NV], ri[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) { bi[v] = 0.0; ri[v] = 0.0; }
  if (i < N) {
    const double mdt = (dtg ? dtg[0] : dt[i]) / dm.vol[i];
    const uint32_t info = dm.info[i];
    // right-hand side (SolverImplicit.cpp:441-473) from the second-order face
    // fluxes of k_face_flux, faces in reference order
#pragma unroll
    for (int q = 0; q < K; q++) {
      const int m = (info >> (2 * q)) & 3;
      const bool own = (info >> (8 + m)) & 1;
      const int fdv = dm.cfd[(size_t)m * N + i];
      const double Af = dm.af[(size_t)m * N + i];
      const double mul = mdt * Af;
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const double F = FF[(size_t)fdv * NV + v];
        if (own) { bi[v] -= mul * F; ri[v] += F * Af; }
        else { bi[v] += mdt * Af * F; ri[v] -= F * Af; }
      }
    }
#pragma unroll
    for (int v = 0; v < NV; v++) b[(size_t)i * NV + v] = bi[v];
    // blocks: identity, then per face (reference order) the repeated adds
    double D[NB];
#pragma unroll
    for (int e = 0; e < NB; e++) D[e] = 0.0;
#pragma unroll
    for (int a = 0; a < NV; a++) D[a * NV + a] += 1.0;
    for (int q = 0; q < K; q++) {
      const int m = (info >> (2 * q)) & 3;
      const bool own = (info >> (8 + m)) & 1;
      const int j = dm.nbr[(size_t)m * N + i];
      const int f = fd.cf[(size_t)m * N + i];
      const double Af = dm.af[(size_t)m * N + i];
      const double mul = own ? mdt * Af : -(mdt * Af);
      // own row: diag += mul*JL, off(i,nbr) = mul*JR; neighbour row:
      // diag += (-m)*JR, off(i,owner) = (-m)*JL
      const size_t FF_ = fd.F;
      const double *Jd = JF + (size_t)(own ? 0 : 1) * NB * FF_ + f;   // entry e at Jd[e * F]
      const double *Jo = JF + (size_t)(own ? 1 : 0) * NB * FF_ + f;
#pragma unroll
      for (int e = 0; e < NB; e++) {
        const double x = mul * Jd[(size_t)e * FF_];
        for (int r = 0; r < repeat; r++) D[e] += x;
      }
      TM *Ao = A + (size_t)(1 + m) * NB * N;
      if (j >= 0) {
#pragma unroll
        for (int e = 0; e < NB; e++) {
          const double x = mul * Jo[(size_t)e * FF_];
          double s = 0.0;
          for (int r = 0; r < repeat; r++) s += x;
          Ao[(size_t)e * N + i] = (TM)s;
        }
      } else {
#pragma unroll
        for (int e = 0; e < NB; e++) Ao[(size_t)e * N + i] = (TM)0;
      }
    }
#pragma unroll
    for (int e = 0; e < NB; e++) A[(size_t)e * N + i] = (TM)D[e];
  }
  if (linf_b) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      double a = block_max(::fabs(bi[v]));
      if (threadIdx.x == 0) atomicMax(&linf_b[v], (unsigned long long)__double_as_longlong(a));
      double c = block_max(::fabs(ri[v]));
      if (threadIdx.x == 0) atomicMax(&linf_r[v], (unsigned long long)__double_as_longlong(c));
    }
  }
}

// block Jacobi: LU slot 0 <- inverse of the diagonal block
template <int ND, typename TM>
__global__ __launch_bounds__(256) void k_bj_setup(const TM *__restrict__ A, TM *__restrict__ LU, int N) {
  constexpr int NV = ND + 2, NB = NV * NV;
  const int i = cell_index();
  if (i >= N) return;
  double D[NB], Inv[NB];
  load_block<TM, NB>(A, 0, N, i, D);
  block_inverse<NV>(D, Inv);
  store_block<TM, NB>(LU, 0, N, i, Inv);
}

// block ILU(0) factorisation of the rows of one colour (all lower colours done).
// LU holds a copy of A on entry; on exit slot 0 holds inv(U_ii), lower slots L_ik,
// upper slots U_ij.
// colour-major copies of the matrix: Ac[slot][e][p] = LU[slot][e][p] = A[slot][e][crow[p]]
template <typename TM>
__global__ void k_ilu_gather(const TM *__restrict__ A, const int *__restrict__ crow, int N, int nslot_e,
                             TM *__restrict__ Ac, TM *__restrict__ LU) {
  const size_t tot = (size_t)nslot_e * N;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x) {
    const size_t se = q / N, p = q - se * N;
    const TM v = A[se * N + crow[p]];
    Ac[q] = v;
    LU[q] = v;
  }
}

// y = A x with matrix and vectors in colour order (neighbour positions cnp)
template <int ND, typename TM, typename TV>
__global__ __launch_bounds__(256) void k_spmv_pos(const TM *__restrict__ A, const int *__restrict__ cnp, int N,
                                                  const TV *__restrict__ x, TV *__restrict__ y) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  double acc[NV], xv[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) xv[c] = (double)x[(size_t)p * NV + c];
#pragma unroll
  for (int r = 0; r < NV; r++) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < NV; c++) s += (double)A[((size_t)(r * NV + c)) * N + p] * xv[c];
    acc[r] = s;
  }
#pragma unroll
  for (int m = 0; m < K; m++) {
    const int q = cnp[(size_t)m * N + p];
    if (q < 0) continue;
#pragma unroll
    for (int c = 0; c < NV; c++) xv[c] = (double)x[(size_t)q * NV + c];
    const TM *Ao = A + (size_t)(1 + m) * NB * N;
#pragma unroll
    for (int r = 0; r < NV; r++) {
      double s = 0;
#pragma unroll
      for (int c = 0; c < NV; c++) s += (double)Ao[((size_t)(r * NV + c)) * N + p] * xv[c];
      acc[r] += s;
    }
  }
#pragma unroll
  for (int r = 0; r < NV; r++) y[(size_t)p * NV + r] = (TV)acc[r];
}

// rows between cell order and colour order: to_colour: dst[p] = src[crow[p]], else dst[crow[p]] = src[p]
template <typename T>
__global__ void k_perm_rows(const T *__restrict__ src, const int *__restrict__ crow, int N, int nv, int to_colour,
                            T *__restrict__ dst) {
  const size_t tot = (size_t)N * nv;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x) {
    const size_t p = q / nv, c = q - p * nv;
    const size_t i = (size_t)crow[p];
    if (to_colour) dst[q] = src[i * nv + c];
    else dst[i * nv + c] = src[q];
  }
}

// block ILU(0) of the rows at positions p0 .. p0+nrows-1 (one colour); rows of
// earlier colours are final.  Elimination order: colour, then cell index (not
// position, so the factors do not depend on how rows are laid out in a colour).
template <int ND, typename TM>
__global__ __launch_bounds__(256) void k_ilu_factor_color(TM *__restrict__ LU, int N, const int *__restrict__ crow,
                                                          const int *__restrict__ cnb, const int *__restrict__ cnp,
                                                          const unsigned char *__restrict__ clow,
                                                          const int *__restrict__ color, TM *__restrict__ Dr, int p0,
                                                          int nrows) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows) return;
  const int p = p0 + t;
  const int i = crow[p];
  const unsigned lowm = clow[p];
  int nb[K], np_[K];
#pragma unroll
  for (int m = 0; m < K; m++) {
    nb[m] = cnb[(size_t)m * N + p];
    np_[m] = cnp[(size_t)m * N + p];
  }
  // lower neighbours in elimination order (colour, then cell index)
  int ord[K], okey[K];
  int nl = 0;
#pragma unroll
  for (int m = 0; m < K; m++)
    if ((lowm >> m) & 1) {
      okey[nl] = color[nb[m]];
      ord[nl++] = m;
    }
  for (int a = 1; a < nl; a++)
    for (int bq = a; bq > 0; bq--) {
      const int x = ord[bq - 1], y = ord[bq];
      const int kx = okey[bq - 1], ky = okey[bq];
      if (kx > ky || (kx == ky && nb[x] > nb[y])) {
        ord[bq - 1] = y; ord[bq] = x;
        okey[bq - 1] = ky; okey[bq] = kx;
      }
    }
  double Aii[NB];
  load_block<TM, NB>(LU, 0, N, p, Aii);
  for (int q = 0; q < nl; q++) {
    const int m = ord[q];
    const int pk = np_[m];
    double Lik[NB], Uinv[NB], Aik[NB];
    load_block<TM, NB>(LU, 1 + m, N, p, Aik);
    // row k is final: inv(U_kk), from the row-major copy (one contiguous block)
#pragma unroll
    for (int e = 0; e < NB; e++) Uinv[e] = (double)Dr[(size_t)pk * NB + e];
#pragma unroll
    for (int r = 0; r < NV; r++)
#pragma unroll
      for (int c = 0; c < NV; c++) {
        double s = 0;
#pragma unroll
        for (int pp = 0; pp < NV; pp++) s += Aik[r * NV + pp] * Uinv[pp * NV + c];
        Lik[r * NV + c] = s;
      }
    store_block<TM, NB>(LU, 1 + m, N, p, Lik);
    // A_ij -= L_ik U_kj for j in row k's upper part that are in row i's pattern
    const unsigned lowk = clow[pk];
#pragma unroll
    for (int mk = 0; mk < K; mk++) {
      const int jk = cnb[(size_t)mk * N + pk];
      if (jk < 0 || ((lowk >> mk) & 1)) continue;   // not upper in row k
      int target = -1;   // -2 = diagonal of row i
      if (jk == i) target = -2;
      else {
#pragma unroll
        for (int mm = 0; mm < K; mm++) if (nb[mm] == jk) target = mm;
      }
      if (target == -1) continue;
      double Ukj[NB];
      load_block<TM, NB>(LU, 1 + mk, N, pk, Ukj);
      double upd[NB];
#pragma unroll
      for (int r = 0; r < NV; r++)
#pragma unroll
        for (int c = 0; c < NV; c++) {
          double s = 0;
#pragma unroll
          for (int pp = 0; pp < NV; pp++) s += Lik[r * NV + pp] * Ukj[pp * NV + c];
          upd[r * NV + c] = s;
        }
      if (target == -2) {
#pragma unroll
        for (int e = 0; e < NB; e++) Aii[e] -= upd[e];
      } else {
        double Aij[NB];
        load_block<TM, NB>(LU, 1 + target, N, p, Aij);
#pragma unroll
        for (int e = 0; e < NB; e++) Aij[e] -= upd[e];
        store_block<TM, NB>(LU, 1 + target, N, p, Aij);
      }
    }
  }
  double Inv[NB];
  block_inverse<NV>(Aii, Inv);
  store_block<TM, NB>(LU, 0, N, p, Inv);
#pragma unroll
  for (int e = 0; e < NB; e++) Dr[(size_t)p * NB + e] = (TM)Inv[e];
}

// forward sweep of one colour (vectors in colour order): y_p = r_p - sum_(k lower) L_pk y_k.
// Rows of colour 0 have no lower neighbours, so y = r there and is never stored:
// positions below y0end read r. With FUSE (the last colour, which has no upper
// neighbours) the backward step x_p = inv(U_pp) y_p follows at once, y_p rounded
// to TV as a store and reload would round it.
template <int ND, typename TM, typename TV, bool FUSE>
__global__ __launch_bounds__(256) void k_ilu_lower(const TM *__restrict__ LU, int N, const int *__restrict__ cnp,
                                                   const unsigned char *__restrict__ clow, int p0, int nrows,
                                                   const TV *__restrict__ r, TV *__restrict__ y, int y0end,
                                                   TV *__restrict__ x) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows) return;
  const int p = p0 + t;
  const unsigned lowm = clow[p];
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) acc[c] = (double)r[(size_t)p * NV + c];
#pragma unroll
  for (int m = 0; m < K; m++) {
    if (!((lowm >> m) & 1)) continue;
    const int q = cnp[(size_t)m * N + p];
    const TV *src = q < y0end ? r : y;
    double yk[NV];
#pragma unroll
    for (int c = 0; c < NV; c++) yk[c] = (double)src[(size_t)q * NV + c];
    const TM *L = LU + (size_t)(1 + m) * NB * N;
#pragma unroll
    for (int a = 0; a < NV; a++) {
      double s = 0;
#pragma unroll
      for (int c = 0; c < NV; c++) s += (double)L[((size_t)(a * NV + c)) * N + p] * yk[c];
      acc[a] -= s;
    }
  }
  if constexpr (FUSE) {
#pragma unroll
    for (int c = 0; c < NV; c++) acc[c] = (double)(TV)acc[c];
#pragma unroll
    for (int a = 0; a < NV; a++) {
      double s = 0;
#pragma unroll
      for (int c = 0; c < NV; c++) s += (double)LU[((size_t)(a * NV + c)) * N + p] * acc[c];
      x[(size_t)p * NV + a] = (TV)s;
    }
  } else {
#pragma unroll
    for (int c = 0; c < NV; c++) y[(size_t)p * NV + c] = (TV)acc[c];
  }
}

// backward sweep of one colour (vectors in colour order): x_p = inv(U_pp) (y_p - sum_(j upper) U_pj x_j);
// y_p is r_p below y0end (colour 0)
template <int ND, typename TM, typename TV>
__global__ __launch_bounds__(256) void k_ilu_upper(const TM *__restrict__ LU, int N, const int *__restrict__ cnp,
                                                   const unsigned char *__restrict__ clow, int p0, int nrows,
                                                   const TV *__restrict__ y, TV *__restrict__ x,
                                                   const TV *__restrict__ r, int y0end) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows) return;
  const int p = p0 + t;
  const unsigned lowm = clow[p];
  const TV *own = p < y0end ? r : y;
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) acc[c] = (double)own[(size_t)p * NV + c];
#pragma unroll
  for (int m = 0; m < K; m++) {
    const int q = cnp[(size_t)m * N + p];
    if (q < 0 || ((lowm >> m) & 1)) continue;
    double xj[NV];
#pragma unroll
    for (int c = 0; c < NV; c++) xj[c] = (double)x[(size_t)q * NV + c];
    const TM *Uo = LU + (size_t)(1 + m) * NB * N;
#pragma unroll
    for (int a = 0; a < NV; a++) {
      double s = 0;
#pragma unroll
      for (int c = 0; c < NV; c++) s += (double)Uo[((size_t)(a * NV + c)) * N + p] * xj[c];
      acc[a] -= s;
    }
  }
#pragma unroll
  for (int a = 0; a < NV; a++) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < NV; c++) s += (double)LU[((size_t)(a * NV + c)) * N + p] * acc[c];
    x[(size_t)p * NV + a] = (TV)s;
  }
}

// ---------------------------------------------------------------------------
// Krylov kernels (vectors [N][nv], type TV)
// ---------------------------------------------------------------------------
// z = inv(D_i) r_i (block Jacobi; also the diagonal step of the U solve)
template <int ND, typename TM, typename TV>
__global__ __launch_bounds__(256) void k_bj_apply(const TM *__restrict__ LU, int N, const TV *__restrict__ r,
                                                  TV *__restrict__ z) {
  constexpr int NV = ND + 2;
  const int i = cell_index();
  if (i >= N) return;
  double ri[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) ri[c] = (double)r[(size_t)i * NV + c];
#pragma unroll
  for (int a = 0; a < NV; a++) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < NV; c++) s += (double)LU[((size_t)(a * NV + c)) * N + i] * ri[c];
    z[(size_t)i * NV + a] = (TV)s;
  }
}

// partial dot products of w with up to 8 basis vectors V[i0 .. i0+cnt), plus ||w||^2
// 16-byte packets of Krylov vectors (4 floats or 2 doubles); vectors start on
// 64-element boundaries (Impl::vstride), so packet loads are aligned
template <typename T>
struct Pk {
  static constexpr int W = 16 / sizeof(T);
  T v[16 / sizeof(T)];
};
template <typename T>
__device__ __forceinline__ Pk<T> ldp(const T *p) {
  Pk<T> r;
  const uint4 u = *reinterpret_cast<const uint4 *>(p);
  __builtin_memcpy(r.v, &u, 16);
  return r;
}
template <typename T>
__device__ __forceinline__ void stp(T *p, const Pk<T> &r) {
  uint4 u;
  __builtin_memcpy(&u, r.v, 16);
  *reinterpret_cast<uint4 *>(p) = u;
}

__device__ __forceinline__ void block_store_sums(const double *acc, int nq, double *part) {
  __shared__ double sh[9][4];
  for (int q = 0; q < nq; q++) {
    const double s = wave_sum(acc[q]);
    if ((threadIdx.x & 63) == 0) sh[q][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if ((int)threadIdx.x < nq) {
    const int q = threadIdx.x;
    part[(size_t)q * gridDim.x + blockIdx.x] = sh[q][0] + sh[q][1] + sh[q][2] + sh[q][3];
  }
}

// up to 8 dot products <V_(i0+q), w> and ||w||^2, per-block partials
template <typename TV>
__global__ __launch_bounds__(256) void k_mdot(const TV *__restrict__ V, size_t stride, int i0, int cnt,
                                              const TV *__restrict__ w, size_t n, double *__restrict__ part) {
  constexpr int W = Pk<TV>::W;
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; q++) acc[q] = 0;
  const size_t np = n / W, tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x;
  for (size_t e = tid; e < np; e += nt) {
    const Pk<TV> wv = ldp(w + e * W);
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (q < cnt) {
        const Pk<TV> vv = ldp(V + (size_t)(i0 + q) * stride + e * W);
#pragma unroll
        for (int t = 0; t < W; t++) acc[q] += (double)vv.v[t] * (double)wv.v[t];
      }
#pragma unroll
    for (int t = 0; t < W; t++) acc[8] += (double)wv.v[t] * (double)wv.v[t];
  }
  for (size_t e = np * W + tid; e < n; e += nt) {   // tail
    const double we = (double)w[e];
    for (int q = 0; q < cnt; q++) acc[q] += (double)V[(size_t)(i0 + q) * stride + e] * we;
    acc[8] += we * we;
  }
  block_store_sums(acc, 9, part);
}

// w -= sum_q h_q V_q (h read from the multidot chunks: h_q = dots[9*(q/8) + q%8]) and ||w||^2
template <typename TV>
__global__ __launch_bounds__(256) void k_orth_update(const TV *__restrict__ V, size_t stride, int cnt,
                                                     const double *__restrict__ dots, TV *__restrict__ w, size_t n,
                                                     double *__restrict__ part) {
  constexpr int W = Pk<TV>::W;
  __shared__ double hs[32];
  if ((int)threadIdx.x < cnt) hs[threadIdx.x] = dots[9 * (threadIdx.x >> 3) + (threadIdx.x & 7)];
  __syncthreads();
  double nrm = 0;
  const size_t np = n / W, tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x;
  for (size_t e = tid; e < np; e += nt) {
    Pk<TV> wv = ldp(w + e * W);
    double we[W];
#pragma unroll
    for (int t = 0; t < W; t++) we[t] = (double)wv.v[t];
    for (int q = 0; q < cnt; q++) {
      const Pk<TV> vv = ldp(V + (size_t)q * stride + e * W);
      const double hq = hs[q];
#pragma unroll
      for (int t = 0; t < W; t++) we[t] -= hq * (double)vv.v[t];
    }
#pragma unroll
    for (int t = 0; t < W; t++) {
      wv.v[t] = (TV)we[t];
      const double wr = (double)wv.v[t];
      nrm += wr * wr;
    }
    stp(w + e * W, wv);
  }
  for (size_t e = np * W + tid; e < n; e += nt) {
    double we = (double)w[e];
    for (int q = 0; q < cnt; q++) we -= hs[q] * (double)V[(size_t)q * stride + e];
    w[e] = (TV)we;
    const double wr = (double)w[e];
    nrm += wr * wr;
  }
  block_store_sums(&nrm, 1, part);
}

// out[q] = sum over blocks of part[q][.], fixed order (deterministic)
__global__ void k_reduce_parts(const double *__restrict__ part, int nblk, int nq, double *__restrict__ out) {
  const int q = blockIdx.x;
  if (q >= nq) return;
  double s = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[(size_t)q * nblk + b];
  __shared__ double sh[4];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[q] = sh[0] + sh[1] + sh[2] + sh[3];
}

// y = a x for two vectors of one type, packets
template <typename TV>
__global__ void k_scale_p(const TV *__restrict__ x, double a, TV *__restrict__ y, size_t n) {
  constexpr int W = Pk<TV>::W;
  const size_t np = n / W, tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x;
  for (size_t e = tid; e < np; e += nt) {
    Pk<TV> v = ldp(x + e * W);
#pragma unroll
    for (int t = 0; t < W; t++) v.v[t] = (TV)(a * (double)v.v[t]);
    stp(y + e * W, v);
  }
  for (size_t e = np * W + tid; e < n; e += nt) y[e] = (TV)(a * (double)x[e]);
}

// w = w / sqrt(n2[0]) when n2[0] > 0 (the norm read on the device, so the
// launch need not wait for the host); same arithmetic as k_scale_p with
// a = 1 / sqrt(n2[0])
template <typename TV>
__global__ void k_scale_norm(TV *__restrict__ w, const double *__restrict__ n2, size_t n) {
  const double s2 = n2[0];
  if (!(s2 > 0)) return;
  const double a = 1.0 / sqrt(s2);
  constexpr int W = Pk<TV>::W;
  const size_t np = n / W, tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x;
  for (size_t e = tid; e < np; e += nt) {
    Pk<TV> v = ldp(w + e * W);
#pragma unroll
    for (int t = 0; t < W; t++) v.v[t] = (TV)(a * (double)v.v[t]);
    stp(w + e * W, v);
  }
  for (size_t e = np * W + tid; e < n; e += nt) w[e] = (TV)(a * (double)w[e]);
}

// y = a*x (scale/copy with conversion)
template <typename TA, typename TB>
__global__ void k_scale(const TA *__restrict__ x, double a, TB *__restrict__ y, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    y[e] = (TB)(a * (double)x[e]);
}

// u = sum_q c[q] V[q]
template <typename TV>
__global__ void k_combine(const TV *__restrict__ V, size_t stride, int cnt, const double *__restrict__ c,
                          TV *__restrict__ u, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    double s = 0;
    for (int q = 0; q < cnt; q++) s += c[q] * (double)V[(size_t)q * stride + e];
    u[e] = (TV)s;
  }
}

// x (FP64) += z ; r = b - y (residual with the FP64 right-hand side), partial ||r||^2
template <typename TV>
__global__ void k_axpy_x(double *__restrict__ x, const TV *__restrict__ z, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    x[e] += (double)z[e];
}
template <typename TV>
__global__ __launch_bounds__(256) void k_resid(const double *__restrict__ b, const TV *__restrict__ y,
                                               TV *__restrict__ r, size_t n, double *__restrict__ part) {
  double nrm = 0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const double v = b[e] - (double)y[e];
    r[e] = (TV)v;
    nrm += v * v;
  }
  double s = wave_sum(nrm);
  __shared__ double sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// implicit update U[v][i] += x[i][v]
template <int ND>
__global__ __launch_bounds__(256) void k_update(double *__restrict__ U, const double *__restrict__ x, int N) {
  constexpr int NV = ND + 2;
  const int i = cell_index();
  if (i >= N) return;
#pragma unroll
  for (int v = 0; v < NV; v++) U[(size_t)v * N + i] += x[(size_t)i * NV + v];
}

__global__ void k_fill_u64_i(unsigned long long *p, int n, unsigned long long v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// BiCGStab vector kernels (AMGCL's bicgstab.hpp, right preconditioning); each
// writes per-block partial sums of the dot products its caller needs next
// ---------------------------------------------------------------------------
// p = r + (-beta*omega) v + beta p   (axpbypcz(1, r, -beta*omega, v, beta, p))
template <typename TV>
__global__ void k_bicg_p(const TV *__restrict__ r, const TV *__restrict__ v, TV *__restrict__ p, double bo,
                         double beta, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    p[e] = (TV)((double)r[e] + bo * (double)v[e] + beta * (double)p[e]);
}
// dots of a with up to two vectors: (a, u), (a, w); per-block partials [2][grid]
template <typename TV>
__global__ __launch_bounds__(256) void k_dot2(const TV *__restrict__ a, const TV *__restrict__ u,
                                              const TV *__restrict__ w, size_t n, double *__restrict__ part) {
  double acc[2] = {0, 0};
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const double ae = (double)a[e];
    acc[0] += ae * (double)u[e];
    if (w) acc[1] += ae * (double)w[e];
  }
  block_store_sums(acc, 2, part);
}
// x += c T (x FP64); o = i + (-c) y; partials (o, o) and (o, h) (h may be null)
template <typename TV>
__global__ __launch_bounds__(256) void k_bicg_step(double *__restrict__ x, const TV *__restrict__ T, double cf,
                                                   const TV *__restrict__ in, const TV *__restrict__ y,
                                                   TV *__restrict__ o, const TV *__restrict__ h, size_t n,
                                                   double *__restrict__ part) {
  double acc[2] = {0, 0};
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    x[e] += cf * (double)T[e];
    const TV oe = (TV)((double)in[e] + (-cf) * (double)y[e]);
    o[e] = oe;
    acc[0] += (double)oe * (double)oe;
    if (h) acc[1] += (double)oe * (double)h[e];
  }
  block_store_sums(acc, 2, part);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static Impl *impl_of(bes_ctx *c) { return (Impl *)c->impl; }

static int impl_alloc(bes_ctx *c, Impl *im) {
  const Geometry &g = c->geo;
  const int N = c->N, k = c->k, nd = c->nd, nv = c->nv, NB = nv * nv;
  const int F = (int)g.nface;
  // faces in device order (sorted by owner cell, as the flux pass) with device
  // cell ids: a cell's Jacobian blocks sit next to its neighbours' in JF
  std::vector<int> fo(F), fn(F), cf((size_t)k * N), mrev((size_t)k * N, -1);
  std::vector<double> fnrm((size_t)nd * F);
  std::vector<int> ref2dev(F);
  for (int q = 0; q < F; q++) ref2dev[c->fref[q]] = q;
  for (int q = 0; q < F; q++) {
    const int64_t f = c->fref[q];
    fo[q] = (int)c->iperm[g.owner[f]];
    if (g.nbr[f] >= 0) fn[q] = (int)c->iperm[g.nbr[f]];
    else fn[q] = (g.btype[f] == kSlipWall) ? -1 : -2;
    for (int d = 0; d < nd; d++) fnrm[(size_t)d * F + q] = g.n[f * nd + d];
  }
  for (int i = 0; i < N; i++) {
    const int64_t e = c->perm[i];
    for (int m = 0; m < k; m++) {
      const int64_t f = g.cface[e * k + m];
      cf[(size_t)m * N + i] = ref2dev[f];
      const int64_t j = g.cadj[e * k + m];
      if (j >= 0)
        for (int mm = 0; mm < k; mm++)
          if (g.cface[j * k + mm] == f) mrev[(size_t)m * N + i] = mm;
    }
  }
  // greedy colouring of the adjacency graph in device (Morton) order
  std::vector<int> color(N, -1);
  int ncol = 0;
  for (int i = 0; i < N; i++) {
    const int64_t e = c->perm[i];
    unsigned used = 0;
    for (int m = 0; m < k; m++) {
      const int64_t j = g.cadj[e * k + m];
      if (j >= 0 && color[c->iperm[j]] >= 0) used |= 1u << color[c->iperm[j]];
    }
    int col = 0;
    while (used & (1u << col)) col++;
    color[i] = col;
    ncol = std::max(ncol, col + 1);
  }
  // Within a colour, rows are grouped by which of their slots are lower (an
  // earlier colour) and which are boundary faces, Morton order inside a group:
  // the sweeps and the factorisation branch per slot on that mask, so waves of
  // one mask load only the blocks they use. Rows of one colour are never
  // coupled and the factorisation orders a row's eliminations by colour and
  // cell index, so this order changes neither the factors nor any row's
  // arithmetic in the sweeps.
  std::vector<unsigned char> ckey(N, 0);
  for (int i = 0; i < N; i++) {
    const int64_t e = c->perm[i];
    for (int m = 0; m < k; m++) {
      const int64_t j = g.cadj[e * k + m];
      if (j < 0) ckey[i] |= (unsigned char)(1u << (4 + m));
      else if (color[c->iperm[j]] < color[i]) ckey[i] |= (unsigned char)(1u << m);
    }
  }
  std::vector<int> crow;
  crow.reserve(N);
  im->color_start.assign(ncol + 1, 0);
  for (int col = 0; col < ncol; col++) {
    im->color_start[col] = (int)crow.size();
    for (int i = 0; i < N; i++) if (color[i] == col) crow.push_back(i);
    std::stable_sort(crow.begin() + im->color_start[col], crow.end(),
                     [&](int a, int b) { return ckey[a] < ckey[b]; });
  }
  im->color_start[ncol] = (int)crow.size();
  im->ncolor = ncol;
  std::vector<int> cpos(N), cnb((size_t)k * N, -1), cnp((size_t)k * N, -1);
  std::vector<unsigned char> clow(N, 0);
  for (int q = 0; q < N; q++) cpos[crow[q]] = q;
  for (int q = 0; q < N; q++) {
    const int64_t e = c->perm[crow[q]];
    for (int m = 0; m < k; m++) {
      const int64_t j = g.cadj[e * k + m];
      if (j < 0) continue;
      const int jd = (int)c->iperm[j];
      cnb[(size_t)m * N + q] = jd;
      cnp[(size_t)m * N + q] = cpos[jd];
      if (cpos[jd] < q) clow[q] |= (unsigned char)(1u << m);
    }
  }
  im->fd.F = F;
  im->ninternal = 0;
  for (int f = 0; f < F; f++) if (g.nbr[f] >= 0) im->ninternal++;
  auto up = [&](auto **p, const auto &h) -> int {
    using T = typename std::decay<decltype(h)>::type::value_type;
    BES_HIP(hipMalloc((void **)p, std::max<size_t>(h.size(), 1) * sizeof(T)));
    BES_HIP(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return BES_OK;
  };
  int s;
  if ((s = up(&im->fd.fo, fo)) || (s = up(&im->fd.fn, fn)) || (s = up(&im->fd.fnrm, fnrm)) ||
      (s = up(&im->fd.cf, cf)) || (s = up(&im->fd.mrev, mrev)) || (s = up(&im->color, color)) ||
      (s = up(&im->crow, crow)) || (s = up(&im->cpos, cpos)) || (s = up(&im->cnb, cnb)) ||
      (s = up(&im->cnp, cnp)) || (s = up(&im->clow, clow)))
    return s;
  const size_t nvec = (size_t)N * nv;
  const size_t es = im->prec == 32 ? sizeof(float) : sizeof(double);
  const size_t mat = (size_t)(k + 1) * NB * N;
  BES_HIP(hipMalloc((void **)&im->JF, (size_t)F * 2 * NB * sizeof(double)));
  BES_HIP(hipMemset(im->JF, 0, (size_t)F * 2 * NB * sizeof(double)));
  BES_HIP(hipMalloc((void **)&im->b, nvec * sizeof(double)));
  BES_HIP(hipMalloc((void **)&im->x, nvec * sizeof(double)));
  BES_HIP(hipMalloc(&im->A, mat * es));
  BES_HIP(hipMalloc(&im->LU, mat * es));
  BES_HIP(hipMalloc(&im->Ac, mat * es));
  BES_HIP(hipMalloc((void **)&im->bc, nvec * sizeof(double)));
  BES_HIP(hipMalloc((void **)&im->xc, nvec * sizeof(double)));
  im->vstride = (nvec + 63) / 64 * 64;   // 16-byte aligned packet loads on every vector
  // GMRES(m) keeps m + 1 basis vectors; BiCGStab's seven vectors use the same storage
  BES_HIP(hipMalloc(&im->V, (size_t)std::max(im->restart + 1, 7) * im->vstride * es));
  BES_HIP(hipMalloc(&im->Z, 3 * im->vstride * es));
  BES_HIP(hipMalloc((void **)&im->part, 32 * kRedBlocks * sizeof(double)));
  BES_HIP(hipMalloc((void **)&im->dots, 64 * sizeof(double)));
  BES_HIP(hipHostMalloc((void **)&im->hdots, 64 * sizeof(double), hipHostMallocDefault));
  BES_HIP(hipEventCreateWithFlags(&im->ev_dots, hipEventDisableTiming));
  BES_HIP(hipMalloc((void **)&im->bits, 16 * sizeof(unsigned long long)));
  return BES_OK;
}

static void impl_release(Impl *im) {
  lp::lp_release(im);
  amg_release(im);
  void *ptrs[] = {im->fd.fo, im->fd.fn, im->fd.fnrm, im->fd.cf, im->fd.mrev, im->color, im->crow, im->JF,
                  im->b, im->x, im->A, im->LU, im->LM, im->LMh, im->V, im->Z, im->part, im->dots, im->bits, im->Ac,
                  im->bc, im->xc,
                  im->cpos, im->cnb, im->cnp, im->clow, im->Dr};
  for (void *p : ptrs)
    if (p) hipFree(p);
  if (im->hdots) hipHostFree(im->hdots);
  if (im->ev_dots) hipEventDestroy(im->ev_dots);
  delete im;
}

template <typename TM>
static int launch_assemble(bes_ctx *c, Impl *im) {
  const int bs = 256, nb = grid_for(c->N, bs);
  FinfI fi;
  std::memcpy(fi.p, c->finf, sizeof(fi.p));
  const double *dtg = c->dt_local ? nullptr : c->dtg;
  const double nbl = (double)c->nv * c->nv;
  {
    // per face: both states, normal, ids; writes one or two nv x nv blocks
    ProfScope ps(c, kProfJacobian, (double)im->fd.F * (16.0 * c->nv + 8.0 * c->nd + 8.0) +
                                       (double)(im->fd.F + im->ninternal) * nbl * 8.0);
    const int nt = (im->fd.F + 63) / 64 * 64 * 2 * c->nv;   // whole groups of 64 faces
    if (c->nd == 2)
      hipLaunchKernelGGL(k_face_jacobian<2>, dim3(grid_for(nt, bs)), dim3(bs), 0, c->stream, c->U0, c->N, im->fd,
                         c->gc, fi, im->JF);
    else
      hipLaunchKernelGGL(k_face_jacobian<3>, dim3(grid_for(nt, bs)), dim3(bs), 0, c->stream, c->U0, c->N, im->fd,
                         c->gc, fi, im->JF);
    BES_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_fill_u64_i, dim3(1), dim3(64), 0, c->stream, im->bits, 16, 0ull);
  int s = launch_face_flux(c);
  if (s) return s;
  // per cell: dt, volume, face records, the cell's face fluxes and both
  // Jacobian blocks of each face; writes b and k+1 matrix blocks
  ProfScope ps(c, kProfAssemble,
               (double)c->N * (20.0 + c->k * 12.0 + 8.0 * c->nv) + (double)c->dm.F * 8.0 * c->nv +
                   (double)(im->fd.F + im->ninternal) * 2.0 * nbl * 8.0 +
                   (double)c->N * (c->k + 1) * nbl * sizeof(TM));
  if (c->nd == 2)
    hipLaunchKernelGGL((k_assemble<2, TM>), dim3(nb), dim3(bs), 0, c->stream, c->FF, c->dt, dtg, c->dm, im->fd,
                       im->JF, im->repeat_nv, im->b, (TM *)im->A, im->bits, im->bits + 8);
  else
    hipLaunchKernelGGL((k_assemble<3, TM>), dim3(nb), dim3(bs), 0, c->stream, c->FF, c->dt, dtg, c->dm, im->fd,
                       im->JF, im->repeat_nv, im->b, (TM *)im->A, im->bits, im->bits + 8);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

template <typename TM>
static int factor(bes_ctx *c, Impl *im) {
  if (im->precond == 4) return amg_build(c, im);   // times itself
  // a lagged learned operator: kept for lp_rebuild - 1 solves after its build
  // (GMRES still applies the current matrix)
  if (im->precond == 3 && im->lp_current && im->lp_since + 1 < im->lp_rebuild) {
    im->lp_since++;
    return BES_OK;
  }
  const int bs = 256, N = c->N, nv = c->nv;
  const size_t mat = (size_t)(c->k + 1) * nv * nv * N;
  const double used = (double)(N + 2 * im->ninternal) * nv * nv * sizeof(TM);   // stored blocks
  ProfScope ps(c, kProfFactor, im->precond == 1 ? 2.0 * N * nv * nv * sizeof(TM) : 3.0 * used);
  if (im->precond == 3) {
    const int s = lp::lp_build(c, im);
    if (s) return s;
    im->lp_since = 0;
    im->lp_current = true;
    im->lp_builds++;
    return BES_OK;
  }
  if (im->precond == 1) {
    if (c->nd == 2) hipLaunchKernelGGL((k_bj_setup<2, TM>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const TM *)im->A, (TM *)im->LU, N);
    else hipLaunchKernelGGL((k_bj_setup<3, TM>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const TM *)im->A, (TM *)im->LU, N);
  } else if (im->precond == 2) {
    if (!im->Dr) BES_HIP(hipMalloc(&im->Dr, (size_t)N * nv * nv * sizeof(TM)));
    hipLaunchKernelGGL((k_ilu_gather<TM>), dim3(2048), dim3(256), 0, c->stream, (const TM *)im->A, im->crow, N,
                       (c->k + 1) * nv * nv, (TM *)im->Ac, (TM *)im->LU);
    for (int col = 0; col < im->ncolor; col++) {
      const int r0 = im->color_start[col], nr = im->color_start[col + 1] - r0;
      if (nr == 0) continue;
      if (c->nd == 2)
        hipLaunchKernelGGL((k_ilu_factor_color<2, TM>), dim3(grid_for(nr, bs)), dim3(bs), 0, c->stream, (TM *)im->LU,
                           N, im->crow, im->cnb, im->cnp, im->clow, im->color, (TM *)im->Dr, r0, nr);
      else
        hipLaunchKernelGGL((k_ilu_factor_color<3, TM>), dim3(grid_for(nr, bs)), dim3(bs), 0, c->stream, (TM *)im->LU,
                           N, im->crow, im->cnb, im->cnp, im->clow, im->color, (TM *)im->Dr, r0, nr);
    }
  }
  BES_HIP(hipGetLastError());
  return BES_OK;
}

// z = M^-1 r
template <typename TM, typename TV>
static int precond_apply(bes_ctx *c, Impl *im, const TV *r, TV *z, TV *tmp) {
  if (im->precond == 4) return amg_apply<TV>(c, im, r, z, tmp);
  const int bs = 256, N = c->N;
  const size_t n = (size_t)N * c->nv;
  const double vb = (double)n * sizeof(TV), blocks = (double)(N + 2 * im->ninternal) * c->nv * c->nv;
  // ILU: every stored block once over both sweeps, r, z, y written and read for
  // the colours between the first and the last, neighbour ids and colours;
  // Jacobi: diagonal inverse, r, z
  const int nc = im->ncolor;
  const double ymid = nc > 1 ? (double)(im->color_start[nc - 1] - im->color_start[1]) / std::max(N, 1) : 0.0;
  ProfScope ps(c, kProfPrecond,
               im->precond == 2 ? blocks * sizeof(TM) + (2.0 + 2.0 * ymid) * vb + 2.0 * N * (4.0 * c->k + 4.0)
               : im->precond == 3 ? blocks * (im->prec == 32 ? 2.0 : (double)sizeof(TM)) + 2.0 * vb + 4.0 * N * c->k
               : (im->precond == 1 ? (double)N * c->nv * c->nv * sizeof(TM) + 2.0 * vb : 2.0 * vb));
  if (im->precond == 0) {
    hipLaunchKernelGGL((k_scale<TV, TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, r, 1.0, z, n);
  } else if (im->precond == 3 && im->prec == 32) {   // learned operator, FP16 copy: one block SpMV
    if (c->nd == 2) hipLaunchKernelGGL((k_spmv<2, __half, TV>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const __half *)im->LMh, c->dm, r, z);
    else hipLaunchKernelGGL((k_spmv<3, __half, TV>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const __half *)im->LMh, c->dm, r, z);
  } else if (im->precond == 3) {   // learned operator: one block SpMV
    if (c->nd == 2) hipLaunchKernelGGL((k_spmv<2, TM, TV>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const TM *)im->LM, c->dm, r, z);
    else hipLaunchKernelGGL((k_spmv<3, TM, TV>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const TM *)im->LM, c->dm, r, z);
  } else if (im->precond == 1) {
    if (c->nd == 2) hipLaunchKernelGGL((k_bj_apply<2, TM, TV>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const TM *)im->LU, N, r, z);
    else hipLaunchKernelGGL((k_bj_apply<3, TM, TV>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const TM *)im->LU, N, r, z);
  } else {
    // colour 0 needs no forward sweep (y = r); the last colour's backward step
    // is fused into its forward sweep
    const int y0end = nc > 0 ? im->color_start[1] : 0;
    for (int col = 1; col < nc; col++) {
      const int r0 = im->color_start[col], nr = im->color_start[col + 1] - r0;
      if (nr == 0) continue;
      const dim3 g(grid_for(nr, bs));
      if (col == nc - 1) {
        if (c->nd == 2)
          hipLaunchKernelGGL((k_ilu_lower<2, TM, TV, true>), g, dim3(bs), 0, c->stream, (const TM *)im->LU, N, im->cnp,
                             im->clow, r0, nr, r, tmp, y0end, z);
        else
          hipLaunchKernelGGL((k_ilu_lower<3, TM, TV, true>), g, dim3(bs), 0, c->stream, (const TM *)im->LU, N, im->cnp,
                             im->clow, r0, nr, r, tmp, y0end, z);
      } else if (c->nd == 2) {
        hipLaunchKernelGGL((k_ilu_lower<2, TM, TV, false>), g, dim3(bs), 0, c->stream, (const TM *)im->LU, N, im->cnp,
                           im->clow, r0, nr, r, tmp, y0end, (TV *)nullptr);
      } else {
        hipLaunchKernelGGL((k_ilu_lower<3, TM, TV, false>), g, dim3(bs), 0, c->stream, (const TM *)im->LU, N, im->cnp,
                           im->clow, r0, nr, r, tmp, y0end, (TV *)nullptr);
      }
    }
    for (int col = nc > 1 ? nc - 2 : nc - 1; col >= 0; col--) {
      const int r0 = im->color_start[col], nr = im->color_start[col + 1] - r0;
      if (nr == 0) continue;
      if (c->nd == 2)
        hipLaunchKernelGGL((k_ilu_upper<2, TM, TV>), dim3(grid_for(nr, bs)), dim3(bs), 0, c->stream, (const TM *)im->LU,
                           N, im->cnp, im->clow, r0, nr, tmp, z, r, y0end);
      else
        hipLaunchKernelGGL((k_ilu_upper<3, TM, TV>), dim3(grid_for(nr, bs)), dim3(bs), 0, c->stream, (const TM *)im->LU,
                           N, im->cnp, im->clow, r0, nr, tmp, z, r, y0end);
    }
  }
  BES_HIP(hipGetLastError());
  return BES_OK;
}

template <typename TM, typename TV>
static int spmv(bes_ctx *c, Impl *im, const TV *x, TV *y) {
  const int bs = 256;
  // stored blocks, x and y once, neighbour ids
  ProfScope ps(c, kProfSpmv, (double)(c->N + 2 * im->ninternal) * c->nv * c->nv * sizeof(TM) +
                                 2.0 * c->N * c->nv * sizeof(TV) + 4.0 * c->N * c->k);
  if (im->precond == 2) {   // ILU(0) solves run in colour order
    if (c->nd == 2)
      hipLaunchKernelGGL((k_spmv_pos<2, TM, TV>), dim3(grid_for(c->N, bs)), dim3(bs), 0, c->stream, (const TM *)im->Ac,
                         im->cnp, c->N, x, y);
    else
      hipLaunchKernelGGL((k_spmv_pos<3, TM, TV>), dim3(grid_for(c->N, bs)), dim3(bs), 0, c->stream, (const TM *)im->Ac,
                         im->cnp, c->N, x, y);
  } else if (c->nd == 2) {
    hipLaunchKernelGGL((k_spmv<2, TM, TV>), dim3(grid_for(c->N, bs)), dim3(bs), 0, c->stream, (const TM *)im->A, c->dm, x, y);
  } else {
    hipLaunchKernelGGL((k_spmv<3, TM, TV>), dim3(grid_for(c->N, bs)), dim3(bs), 0, c->stream, (const TM *)im->A, c->dm, x, y);
  }
  BES_HIP(hipGetLastError());
  return BES_OK;
}

// dots h[0..cnt) = V[0..cnt)^T w and ||w||^2 (last), copied to host; with
// fetch false only queued (chunk q in dots[9q .. 9q+8], ||w||^2 in dots[8])
template <typename TV>
static int multidot(bes_ctx *c, Impl *im, const TV *V, size_t stride, int cnt, const TV *w, size_t n, double *h,
                    double *wnorm2, bool fetch = true) {
  // chunk q (8 dots + ||w||^2) reduces into dots[9q .. 9q+8]; at most 4 chunks
  const int nchunk = std::max(1, (cnt + 7) / 8);
  if (nchunk > 4) { set_error("too many basis vectors"); return BES_ERR_ARG; }
  {
    ProfScope ps(c, kProfKrylov, (double)(cnt + nchunk) * n * sizeof(TV));
    for (int q = 0; q < nchunk; q++) {
      const int i0 = q * 8, cc = std::min(8, cnt - i0);
      hipLaunchKernelGGL((k_mdot<TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, V, stride, i0, std::max(cc, 0), w,
                         n, im->part);
      hipLaunchKernelGGL(k_reduce_parts, dim3(9), dim3(256), 0, c->stream, im->part, kRedBlocks, 9, im->dots + 9 * q);
    }
  }
  BES_HIP(hipGetLastError());
  if (!fetch) return BES_OK;
  double tmp[36];
  BES_HIP(hipMemcpyAsync(tmp, im->dots, sizeof(double) * 9 * nchunk, hipMemcpyDeviceToHost, c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  for (int q = 0; q < cnt; q++) h[q] = tmp[9 * (q / 8) + (q % 8)];
  if (wnorm2) *wnorm2 = tmp[8];
  return BES_OK;
}

template <typename TM, typename TV>
static int gmres(bes_ctx *c, Impl *im, int *iters_out, double *relres_out) {
  const int m = im->restart;
  const size_t n = (size_t)c->N * c->nv, vs = im->vstride;
  TV *V = (TV *)im->V;
  TV *Z = (TV *)im->Z, *Z1 = Z + vs, *Z2 = Z + 2 * vs;
  // ILU(0) solves run in colour order: b in, x out are permuted once per solve
  const bool colour = im->precond == 2;
  const double *bb = im->b;
  double *xx = colour ? im->xc : im->x;
  if (colour) {
    hipLaunchKernelGGL(k_perm_rows<double>, dim3(1024), dim3(256), 0, c->stream, bb, im->crow, c->N, c->nv, 1,
                       im->bc);
    bb = im->bc;
  }
  hipMemsetAsync(xx, 0, n * sizeof(double), c->stream);
  // ||b||
  double bnorm2 = 0;
  hipLaunchKernelGGL((k_scale<double, TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, bb, 1.0, Z, n);
  {
    double dummy;
    int s = multidot<TV>(c, im, Z, vs, 0, Z, n, &dummy, &bnorm2);
    if (s) return s;
  }
  const double bnorm = std::sqrt(bnorm2);
  *iters_out = 0;
  *relres_out = 0;
  if (bnorm == 0) {
    if (colour) hipMemsetAsync(im->x, 0, n * sizeof(double), c->stream);
    return BES_OK;
  }
  const double target = im->rtol * bnorm;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), hcol(m + 1);
  int total = 0;
  double rnorm = bnorm;
  bool first = true, est_done = false;
  while (total < im->maxiter) {
    // r = b - A x (x = 0 on the first cycle, as the reference starts every solve).
    // Every later cycle starts from the true residual of the current x, and the
    // solve ends only when that is at the target: the Givens estimate of the
    // cycle before can fall below the true residual (rounding in the basis,
    // FP32 Krylov vectors), as AMGCL's FGMRES also rechecks (fgmres.hpp:158-163)
    if (first) {
      hipLaunchKernelGGL((k_scale<double, TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, bb, 1.0 / bnorm, V, n);
      rnorm = bnorm;
    } else {
      // x in FP64 -> Z2 (TV), y = A Z2, r = b - y
      hipLaunchKernelGGL((k_scale<double, TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, xx, 1.0, Z2, n);
      int s = spmv<TM, TV>(c, im, Z2, Z1);
      if (s) return s;
      hipLaunchKernelGGL((k_resid<TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, bb, Z1, Z, n, im->part);
      hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, c->stream, im->part, kRedBlocks, 1, im->dots);
      double r2;
      BES_HIP(hipMemcpyAsync(&r2, im->dots, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      BES_HIP(hipStreamSynchronize(c->stream));
      rnorm = std::sqrt(r2);
      if (rnorm <= target) break;
      if (est_done) im->extra_restarts++;   // the estimate said done, the true residual did not
      hipLaunchKernelGGL((k_scale_p<TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, Z, 1.0 / rnorm, V, n);
    }
    first = false;
    est_done = false;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = rnorm;
    int j = 0;
    bool done = false, queued = false;
    for (; j < m && total < im->maxiter; j++) {
      TV *vj = V + (size_t)j * vs, *w = V + (size_t)(j + 1) * vs;
      int s;
      if (!queued) {
        if ((s = precond_apply<TM, TV>(c, im, vj, Z, Z1))) return s;
        if ((s = spmv<TM, TV>(c, im, Z, w))) return s;
      }
      queued = false;
      // classical Gram-Schmidt with one DGKS re-orthogonalisation. The dots,
      // the update and (without re-orthogonalisation) the normalisation are
      // queued back to back and the host waits once, for h and ||w||.
      double w0n2 = 0, wn2 = 0;
      const bool dev_scale = !(im->reorth_eta > 0);
      if ((s = multidot<TV>(c, im, V, vs, j + 1, w, n, nullptr, nullptr, false))) return s;
      {
        ProfScope ps(c, kProfKrylov, (double)(j + 3) * n * sizeof(TV));
        hipLaunchKernelGGL((k_orth_update<TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, V, vs, j + 1, im->dots, w,
                           n, im->part);
        hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, c->stream, im->part, kRedBlocks, 1, im->dots + 40);
      }
      if (dev_scale)
        hipLaunchKernelGGL((k_scale_norm<TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, w, im->dots + 40, n);
      BES_HIP(hipMemcpyAsync(im->hdots, im->dots, 41 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      BES_HIP(hipEventRecord(im->ev_dots, c->stream));
      // the next step's preconditioner and product need only the normalised w:
      // queue them before waiting, so the device works while the host reads h
      // and ||w|| (their results are discarded if this step ends the cycle)
      if (dev_scale && j + 1 < m && total + 1 < im->maxiter) {
        if ((s = precond_apply<TM, TV>(c, im, w, Z, Z1))) return s;
        if ((s = spmv<TM, TV>(c, im, Z, V + (size_t)(j + 2) * vs))) return s;
        queued = true;
      }
      BES_HIP(hipEventSynchronize(im->ev_dots));
      for (int q = 0; q <= j; q++) hcol[q] = im->hdots[9 * (q / 8) + (q % 8)];
      w0n2 = im->hdots[8];
      wn2 = im->hdots[40];
      if (!dev_scale && std::sqrt(wn2) < im->reorth_eta * std::sqrt(w0n2)) {
        std::vector<double> h2(j + 1);
        if ((s = multidot<TV>(c, im, V, vs, j + 1, w, n, h2.data(), nullptr))) return s;
        for (int q = 0; q <= j; q++) hcol[q] += h2[q];
        im->reorths++;
        {
          ProfScope ps(c, kProfKrylov, (double)(j + 3) * n * sizeof(TV));
          hipLaunchKernelGGL((k_orth_update<TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, V, vs, j + 1, im->dots,
                             w, n, im->part);
          hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, c->stream, im->part, kRedBlocks, 1, im->dots + 40);
        }
        BES_HIP(hipMemcpyAsync(&wn2, im->dots + 40, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        BES_HIP(hipStreamSynchronize(c->stream));
      }
      const double hn = std::sqrt(wn2);
      hcol[j + 1] = hn;
      if (hn > 0 && !dev_scale)
        hipLaunchKernelGGL((k_scale_p<TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, w, 1.0 / hn, w, n);
      // Givens rotations
      for (int q = 0; q < j; q++) {
        const double t = cs[q] * hcol[q] + sn[q] * hcol[q + 1];
        hcol[q + 1] = -sn[q] * hcol[q] + cs[q] * hcol[q + 1];
        hcol[q] = t;
      }
      const double den = std::hypot(hcol[j], hcol[j + 1]);
      cs[j] = den > 0 ? hcol[j] / den : 1.0;
      sn[j] = den > 0 ? hcol[j + 1] / den : 0.0;
      hcol[j] = den;
      hcol[j + 1] = 0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      for (int q = 0; q <= j; q++) H[(size_t)q * m + j] = hcol[q];
      total++;
      rnorm = std::fabs(g[j + 1]);
      if (rnorm <= target || hn == 0) { j++; done = true; break; }
    }
    // solve H y = g (upper triangular j x j)
    for (int q = j - 1; q >= 0; q--) {
      double s = g[q];
      for (int p = q + 1; p < j; p++) s -= H[(size_t)q * m + p] * y[p];
      y[q] = s / H[(size_t)q * m + q];
    }
    BES_HIP(hipMemcpyAsync(im->dots, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL((k_combine<TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, V, vs, j, im->dots, Z2, n);
    int s;
    if ((s = precond_apply<TM, TV>(c, im, Z2, Z, Z1))) return s;
    hipLaunchKernelGGL((k_axpy_x<TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, xx, Z, n);
    BES_HIP(hipGetLastError());
    est_done = done;   // confirmed (or the solve restarted) on the true residual above
  }
  // true residual of the returned solution
  {
    hipLaunchKernelGGL((k_scale<double, TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, xx, 1.0, Z2, n);
    int s = spmv<TM, TV>(c, im, Z2, Z1);
    if (s) return s;
    hipLaunchKernelGGL((k_resid<TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, bb, Z1, Z, n, im->part);
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, c->stream, im->part, kRedBlocks, 1, im->dots);
    double r2;
    BES_HIP(hipMemcpyAsync(&r2, im->dots, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BES_HIP(hipStreamSynchronize(c->stream));
    rnorm = std::sqrt(r2);
  }
  if (colour)
    hipLaunchKernelGGL(k_perm_rows<double>, dim3(1024), dim3(256), 0, c->stream, im->xc, im->crow, c->N, c->nv, 0,
                       im->x);
  *iters_out = total;
  *relres_out = rnorm / bnorm;
  return BES_OK;
}

// reduce nq per-block partial sums (written by a kRedBlocks-block kernel) and read them on the host
static int fetch_sums(bes_ctx *c, Impl *im, int nq, double *out) {
  hipLaunchKernelGGL(k_reduce_parts, dim3(nq), dim3(256), 0, c->stream, im->part, kRedBlocks, nq, im->dots);
  BES_HIP(hipGetLastError());
  BES_HIP(hipMemcpyAsync(im->hdots, im->dots, nq * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  for (int q = 0; q < nq; q++) out[q] = im->hdots[q];
  return BES_OK;
}

// BiCGStab with right preconditioning, as AMGCL's solver (bicgstab.hpp:
// operator()): x = 0, r = rh = b; per iteration rho = (r, rh), p, T = M p,
// v = A T, alpha = rho / (rh, v), x += alpha T, s = r - alpha v; if ||s|| is
// above the target: T = M s, t = A T, omega = (t, s) / (t, t), x += omega T,
// r = s - omega t.  Stops when the recursive residual norm is at the target or
// after maxiter iterations; returns the true relative residual.
template <typename TM, typename TV>
static int bicgstab(bes_ctx *c, Impl *im, int *iters_out, double *relres_out) {
  const size_t n = (size_t)c->N * c->nv, vs = im->vstride;
  TV *V = (TV *)im->V;
  TV *r = V, *rh = V + vs, *p = V + 2 * vs, *v = V + 3 * vs, *sv = V + 4 * vs, *t = V + 5 * vs, *T = V + 6 * vs;
  TV *Z = (TV *)im->Z, *Z1 = Z + vs, *Z2 = Z + 2 * vs;
  const bool colour = im->precond == 2;
  const double *bb = im->b;
  double *xx = colour ? im->xc : im->x;
  const dim3 G(kRedBlocks), B(256);
  if (colour) {
    hipLaunchKernelGGL(k_perm_rows<double>, dim3(1024), B, 0, c->stream, bb, im->crow, c->N, c->nv, 1, im->bc);
    bb = im->bc;
  }
  hipMemsetAsync(xx, 0, n * sizeof(double), c->stream);
  hipLaunchKernelGGL((k_scale<double, TV>), G, B, 0, c->stream, bb, 1.0, r, n);
  hipLaunchKernelGGL((k_scale<double, TV>), G, B, 0, c->stream, bb, 1.0, rh, n);
  double d[2];
  int st;
  hipLaunchKernelGGL((k_dot2<TV>), G, B, 0, c->stream, r, r, (const TV *)nullptr, n, im->part);
  if ((st = fetch_sums(c, im, 1, d))) return st;
  const double bnorm = std::sqrt(d[0]);
  *iters_out = 0;
  *relres_out = 0;
  if (bnorm == 0) {
    if (colour) hipMemsetAsync(im->x, 0, n * sizeof(double), c->stream);
    return BES_OK;
  }
  const double target = im->rtol * bnorm;
  double res = bnorm, rho1 = d[0], rho2 = 0, alpha = 0, omega = 0;   // (r, rh) = ||b||^2 at the start
  int iter = 0;
  for (bool first = true; res > target && iter < im->maxiter; ++iter) {
    rho2 = rho1;
    if (iter > 0) rho1 = d[1];   // (r, rh) of the r the last iteration ended with
    if (first) {
      hipLaunchKernelGGL((k_scale<TV, TV>), G, B, 0, c->stream, r, 1.0, p, n);
      first = false;
    } else {
      if (rho2 == 0 || omega == 0) { set_error("BiCGStab breakdown (zero rho or omega)"); break; }
      const double beta = (rho1 * alpha) / (rho2 * omega);
      hipLaunchKernelGGL((k_bicg_p<TV>), G, B, 0, c->stream, r, v, p, -beta * omega, beta, n);
    }
    if ((st = precond_apply<TM, TV>(c, im, p, T, Z1))) return st;
    if ((st = spmv<TM, TV>(c, im, T, v))) return st;
    hipLaunchKernelGGL((k_dot2<TV>), G, B, 0, c->stream, rh, v, (const TV *)nullptr, n, im->part);
    if ((st = fetch_sums(c, im, 1, d))) return st;
    if (d[0] == 0) { set_error("BiCGStab breakdown ((rh, v) = 0)"); break; }
    alpha = rho1 / d[0];
    hipLaunchKernelGGL((k_bicg_step<TV>), G, B, 0, c->stream, xx, T, alpha, r, v, sv, (const TV *)nullptr, n,
                       im->part);
    if ((st = fetch_sums(c, im, 1, d))) return st;
    res = std::sqrt(d[0]);
    if (res > target) {
      if ((st = precond_apply<TM, TV>(c, im, sv, T, Z1))) return st;
      if ((st = spmv<TM, TV>(c, im, T, t))) return st;
      hipLaunchKernelGGL((k_dot2<TV>), G, B, 0, c->stream, t, sv, t, n, im->part);
      if ((st = fetch_sums(c, im, 2, d))) return st;
      omega = d[1] > 0 ? d[0] / d[1] : 0.0;
      if (omega == 0) { set_error("BiCGStab breakdown (omega = 0)"); ++iter; break; }
      hipLaunchKernelGGL((k_bicg_step<TV>), G, B, 0, c->stream, xx, T, omega, sv, t, r, rh, n, im->part);
      if ((st = fetch_sums(c, im, 2, d))) return st;
      res = std::sqrt(d[0]);
    }
  }
  // true residual of the returned solution
  hipLaunchKernelGGL((k_scale<double, TV>), G, B, 0, c->stream, xx, 1.0, Z2, n);
  if ((st = spmv<TM, TV>(c, im, Z2, Z1))) return st;
  hipLaunchKernelGGL((k_resid<TV>), G, B, 0, c->stream, bb, Z1, Z, n, im->part);
  if ((st = fetch_sums(c, im, 1, d))) return st;
  if (colour)
    hipLaunchKernelGGL(k_perm_rows<double>, dim3(1024), B, 0, c->stream, im->xc, im->crow, c->N, c->nv, 0, im->x);
  BES_HIP(hipGetLastError());
  *iters_out = iter;
  *relres_out = std::sqrt(d[0]) / bnorm;
  return BES_OK;
}

}  // namespace bes

using namespace bes;

extern "C" {

void bes_implicit_free(bes_ctx *c) {
  if (c && c->impl) {
    impl_release((Impl *)c->impl);
    c->impl = nullptr;
  }
}

int bes_implicit_setup(bes_ctx *c, int precision, int precond, int restart, int maxiter, double rtol) {
  if (!c || (precision != 64 && precision != 32) || precond < 0 || precond > 4 || restart < 1 || restart > 30 ||
      maxiter < 1 || !(rtol > 0)) {
    set_error("invalid implicit setup argument (precision 64/32, precond 0-4, restart 1-30)");
    return BES_ERR_ARG;
  }
  if (c->B != 1) {
    set_error("the implicit path runs one condition (batch size 1)");
    return BES_ERR_ARG;
  }
  bes_implicit_free(c);
  Impl *im = new Impl();
  im->prec = precision;
  im->precond = precond;
  im->restart = restart;
  im->maxiter = maxiter;
  im->rtol = rtol;
  c->impl = im;
  int s = impl_alloc(c, im);
  if (s) { bes_implicit_free(c); return s; }
  return BES_OK;
}

int bes_implicit_set_reorth(bes_ctx *c, double eta) {
  if (!c || !c->impl || !(eta >= 0.0) || eta > 1.0) { set_error("call bes_implicit_setup first / eta in [0,1]"); return BES_ERR_ARG; }
  impl_of(c)->reorth_eta = eta;
  return BES_OK;
}

int bes_implicit_stats(bes_ctx *c, long *reorths, int *lp_fallbacks, long *extra_restarts) {
  if (!c || !c->impl) { set_error("call bes_implicit_setup first"); return BES_ERR_ARG; }
  if (reorths) *reorths = impl_of(c)->reorths;
  if (lp_fallbacks) *lp_fallbacks = impl_of(c)->lp_fallbacks;
  if (extra_restarts) *extra_restarts = impl_of(c)->extra_restarts;
  return BES_OK;
}

int bes_implicit_set_solver(bes_ctx *c, int solver) {
  if (!c || !c->impl) { set_error("call bes_implicit_setup first"); return BES_ERR_ARG; }
  if (solver != BES_SOLVER_GMRES && solver != BES_SOLVER_BICGSTAB) {
    set_error("LINEAR_SOLVER: only GMRES (FGMRES) and BICGSTAB are available; LGMRES is not");
    return BES_ERR_ARG;
  }
  impl_of(c)->solver = solver;
  return BES_OK;
}

int bes_implicit_set_precond(bes_ctx *c, int precond) {
  if (!c || !c->impl || precond < 0 || precond > 4) { set_error("call bes_implicit_setup first / precond 0-4"); return BES_ERR_ARG; }
  Impl *im = impl_of(c);
  im->precond = precond;
  im->factored = false;
  return BES_OK;
}

int bes_implicit_assemble(bes_ctx *c, double cfl, int local, int repeat_nv, double *linf_b, double *linf_r) {
  if (!c || !c->impl) { set_error("call bes_implicit_setup first"); return BES_ERR_ARG; }
  Impl *im = impl_of(c);
  im->repeat_nv = repeat_nv ? c->nv : 1;
  int s;
  hipLaunchKernelGGL(k_fill_u64_i, dim3(1), dim3(64), 0, c->stream, c->err, 1, ~0ull);
  if ((s = launch_timestep(c, c->U0, cfl, local))) return s;
  c->dt_local = local != 0;
  if ((s = launch_gradlim(c, c->U0, nullptr, nullptr))) return s;
  if (im->prec == 64) s = launch_assemble<double>(c, im);
  else s = launch_assemble<float>(c, im);
  if (s) return s;
  unsigned long long bb[16];
  BES_HIP(hipMemcpyAsync(bb, im->bits, sizeof(bb), hipMemcpyDeviceToHost, c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  for (int v = 0; v < c->nv; v++) {
    if (linf_b) std::memcpy(&linf_b[v], &bb[v], 8);
    if (linf_r) std::memcpy(&linf_r[v], &bb[8 + v], 8);
  }
  im->assembled = true;
  im->factored = false;
  unsigned long long e = 0;
  BES_HIP(hipMemcpy(&e, c->err, sizeof(e), hipMemcpyDeviceToHost));
  if (e != ~0ull) {
    c->last_err_face = e < c->fref.size() ? (long)c->fref[e] : -1;   // device face -> reference id
    set_error("non-finite flux at face " + std::to_string(c->last_err_face));
    return BES_ERR_NAN;
  }
  return BES_OK;
}

// reference CoefficientMatrix layout: scalar CSR rows (cell, var) in reference
// cell order, each row = own block row then neighbour block rows in local face
// order (CoefficientMatrix.cpp:127-160).  rowptr [ncell*nv+1], col/val [nnz].
int bes_implicit_get_system(bes_ctx *c, long *rowptr, long *col, double *val, double *rhs) {
  if (!c || !c->impl || !impl_of(c)->assembled) { set_error("system not assembled"); return BES_ERR_ARG; }
  Impl *im = impl_of(c);
  const int N = c->N, nv = c->nv, k = c->k, NB = nv * nv;
  const size_t mat = (size_t)(k + 1) * NB * N;
  std::vector<double> A(mat);
  if (im->prec == 64) {
    BES_HIP(hipMemcpy(A.data(), im->A, mat * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> Af(mat);
    BES_HIP(hipMemcpy(Af.data(), im->A, mat * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < mat; q++) A[q] = Af[q];
  }
  const Geometry &g = c->geo;
  long nz = 0, row = 0;
  for (int64_t e = 0; e < N; e++) {
    const int i = (int)c->iperm[e];
    for (int a = 0; a < nv; a++) {
      if (rowptr) rowptr[row] = nz;
      for (int bq = 0; bq < nv; bq++) {
        if (col) col[nz] = e * nv + bq;
        if (val) val[nz] = A[((size_t)0 * NB + a * nv + bq) * N + i];
        nz++;
      }
      for (int m = 0; m < k; m++) {
        const int64_t j = g.cadj[e * k + m];
        if (j < 0) continue;
        for (int bq = 0; bq < nv; bq++) {
          if (col) col[nz] = j * nv + bq;
          if (val) val[nz] = A[((size_t)(1 + m) * NB + a * nv + bq) * N + i];
          nz++;
        }
      }
      row++;
    }
  }
  if (rowptr) rowptr[row] = nz;
  if (rhs) {
    std::vector<double> b((size_t)N * nv);
    BES_HIP(hipMemcpy(b.data(), im->b, b.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t e = 0; e < N; e++)
      for (int v = 0; v < nv; v++) rhs[e * nv + v] = b[(size_t)c->iperm[e] * nv + v];
  }
  return BES_OK;
}

// load a system assembled on the host in the reference CSR layout (the
// CoefficientMatrix::addToBlock / addToVal path, CoefficientMatrix.cpp:166-204)
int bes_implicit_set_system(bes_ctx *c, const double *val, const double *rhs) {
  if (!c || !c->impl || !val || !rhs) { set_error("call bes_implicit_setup first / null argument"); return BES_ERR_ARG; }
  Impl *im = impl_of(c);
  const int N = c->N, nv = c->nv, k = c->k, NB = nv * nv;
  const size_t mat = (size_t)(k + 1) * NB * N;
  std::vector<double> A(mat, 0.0), b((size_t)N * nv);
  const Geometry &g = c->geo;
  long nz = 0;
  for (int64_t e = 0; e < N; e++) {
    const int i = (int)c->iperm[e];
    for (int a = 0; a < nv; a++) {
      for (int bq = 0; bq < nv; bq++) A[((size_t)a * nv + bq) * N + i] = val[nz++];
      for (int m = 0; m < k; m++) {
        if (g.cadj[e * k + m] < 0) continue;
        for (int bq = 0; bq < nv; bq++) A[((size_t)(1 + m) * NB + a * nv + bq) * N + i] = val[nz++];
      }
    }
    for (int v = 0; v < nv; v++) b[(size_t)i * nv + v] = rhs[e * nv + v];
  }
  if (im->prec == 64) {
    BES_HIP(hipMemcpy(im->A, A.data(), mat * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> Af(A.begin(), A.end());
    BES_HIP(hipMemcpy(im->A, Af.data(), mat * sizeof(float), hipMemcpyHostToDevice));
  }
  BES_HIP(hipMemcpy(im->b, b.data(), b.size() * sizeof(double), hipMemcpyHostToDevice));
  im->assembled = true;
  im->factored = false;
  return BES_OK;
}

long bes_implicit_nnz(bes_ctx *c) {
  if (!c) return -1;
  long nz = 0;
  for (int64_t e = 0; e < c->N; e++) {
    nz++;
    for (int m = 0; m < c->k; m++) if (c->geo.cadj[e * c->k + m] >= 0) nz++;
  }
  return nz * c->nv * c->nv;
}

static int solve_once(bes_ctx *c, Impl *im, int *it, double *rr) {
  int s;
  if (im->prec == 64) {
    if (!im->factored && (s = factor<double>(c, im))) return s;
    im->factored = true;
    return im->solver == 1 ? bicgstab<double, double>(c, im, it, rr) : gmres<double, double>(c, im, it, rr);
  }
  if (!im->factored && (s = factor<float>(c, im))) return s;
  im->factored = true;
  return im->solver == 1 ? bicgstab<float, float>(c, im, it, rr) : gmres<float, float>(c, im, it, rr);
}

int bes_linear_solve(bes_ctx *c, int *iters, double *relres) {
  if (!c || !c->impl || !impl_of(c)->assembled) { set_error("system not assembled"); return BES_ERR_ARG; }
  Impl *im = impl_of(c);
  if (im->precond == 3 && !im->lp) { set_error("learned preconditioner not initialised (bes_lp_init)"); return BES_ERR_ARG; }
  int s;
  int it = 0;
  double rr = 0;
  if ((s = solve_once(c, im, &it, &rr))) return s;
  if (im->precond == 3 && !(rr <= im->rtol * 1.0001) && !(rr == 0)) {
    // the learned operator stalled: solve again with multicolour ILU(0)
    int it2 = 0;
    im->precond = 2;
    im->factored = false;
    s = solve_once(c, im, &it2, &rr);
    im->precond = 3;
    im->factored = false;
    im->lp_current = false;   // rebuild the operator for the next solve
    im->lp_fallbacks++;
    it += it2;
    if (s) return s;
  }
  if ((s = prof_collect(c))) return s;
  im->last_iters = it;
  im->last_relres = rr;
  if (iters) *iters = it;
  if (relres) *relres = rr;
  if (!(rr <= im->rtol * 1.0001) && !(rr == 0)) {
    set_error("linear solver stopped at relative residual " + std::to_string(rr));
    return BES_ERR_SOLVE;
  }
  return BES_OK;
}

int bes_get_linear_solution(bes_ctx *c, double *x) {
  if (!c || !c->impl || !x) { set_error("no solution"); return BES_ERR_ARG; }
  Impl *im = impl_of(c);
  const int N = c->N, nv = c->nv;
  std::vector<double> h((size_t)N * nv);
  BES_HIP(hipMemcpy(h.data(), im->x, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int64_t e = 0; e < N; e++)
    for (int v = 0; v < nv; v++) x[e * nv + v] = h[(size_t)c->iperm[e] * nv + v];
  return BES_OK;
}

int bes_amg_apply(bes_ctx *c, const double *r, double *z) {
  if (!c || !c->impl || !r || !z) { set_error("call bes_implicit_setup first"); return BES_ERR_ARG; }
  Impl *im = impl_of(c);
  if (im->precond != 4 || !im->assembled) { set_error("AMG apply needs precond 4 and an assembled system"); return BES_ERR_ARG; }
  int s;
  if (!im->factored) {
    if ((s = im->prec == 64 ? factor<double>(c, im) : factor<float>(c, im))) return s;
    im->factored = true;
  }
  const int N = c->N, nv = c->nv;
  const size_t n = (size_t)N * nv, vs = im->vstride;
  std::vector<double> h(n);
  for (int64_t e = 0; e < N; e++)
    for (int v = 0; v < nv; v++) h[(size_t)c->iperm[e] * nv + v] = r[e * nv + v];
  if (im->prec == 64) {
    double *Z = (double *)im->Z;
    BES_HIP(hipMemcpyAsync(Z, h.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((s = amg_apply<double>(c, im, Z, Z + vs, Z + 2 * vs))) return s;
    BES_HIP(hipMemcpyAsync(h.data(), Z + vs, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  } else {
    float *Z = (float *)im->Z;
    std::vector<float> hf(h.begin(), h.end());
    BES_HIP(hipMemcpyAsync(Z, hf.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if ((s = amg_apply<float>(c, im, Z, Z + vs, Z + 2 * vs))) return s;
    BES_HIP(hipMemcpyAsync(hf.data(), Z + vs, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    BES_HIP(hipStreamSynchronize(c->stream));
    std::copy(hf.begin(), hf.end(), h.begin());
  }
  BES_HIP(hipStreamSynchronize(c->stream));
  for (int64_t e = 0; e < N; e++)
    for (int v = 0; v < nv; v++) z[e * nv + v] = h[(size_t)c->iperm[e] * nv + v];
  return BES_OK;
}

int bes_implicit_update(bes_ctx *c) {
  if (!c || !c->impl) { set_error("call bes_implicit_setup first"); return BES_ERR_ARG; }
  Impl *im = impl_of(c);
  const int bs = 256;
  {
    ProfScope ps(c, kProfUpdate, 24.0 * c->N * c->nv);
    if (c->nd == 2) hipLaunchKernelGGL(k_update<2>, dim3(grid_for(c->N, bs)), dim3(bs), 0, c->stream, c->U0, im->x, c->N);
    else hipLaunchKernelGGL(k_update<3>, dim3(grid_for(c->N, bs)), dim3(bs), 0, c->stream, c->U0, im->x, c->N);
  }
  BES_HIP(hipGetLastError());
  BES_HIP(hipStreamSynchronize(c->stream));
  return prof_collect(c);
}

// SolverImplicit::solve loop (SolverImplicit.cpp:552-600) with the adaptive
// Courant number of SolverImplicit.cpp:530-550.  hist [nsteps][2*nv+3]:
// Linf(b) per var, Linf(sum F*A) per var, CFL used, Krylov iterations, relres.
// nsteps more pseudo-time steps of the loop in Impl::run
static int run_steps(bes_ctx *c, Impl *im, int nsteps, double *hist, int *steps_done) {
  const int nv = c->nv;
  Impl::Run &R = im->run;
  std::vector<double> lb(nv), lr(nv);
  int done = 0;
  for (int it = 0; it < nsteps && !R.reached; it++) {
    int s = bes_implicit_assemble(c, R.cfl, R.local, R.repeat_nv, lb.data(), lr.data());
    if (s) { if (steps_done) *steps_done = done; return s; }
    int iters = 0;
    double rr = 0;
    s = bes_linear_solve(c, &iters, &rr);
    if (s && s != BES_ERR_SOLVE) { if (steps_done) *steps_done = done; return s; }
    if ((s = bes_implicit_update(c))) { if (steps_done) *steps_done = done; return s; }
    if (hist) {
      double *h = hist + (size_t)it * (2 * nv + 3);
      for (int v = 0; v < nv; v++) { h[v] = lb[v]; h[nv + v] = lr[v]; }
      h[2 * nv] = R.cfl;
      h[2 * nv + 1] = iters;
      h[2 * nv + 2] = rr;
    }
    done = it + 1;
    if (R.r0 < 0) R.r0 = lr[0];
    // adaptive Courant number (density monitor of the implicit residual)
    if (R.cfl_max > R.cfl_min) {
      const int ei = (R.lb_old[0] > lb[0]) ? 1 : 0;
      R.cfl *= std::pow(R.lb_old[0] / lb[0], R.ex[ei]);
      R.cfl = std::max(R.cfl_min, R.cfl);
      R.cfl = std::min(R.cfl_max, R.cfl);
    }
    R.lb_old = lb;
    if (R.stop_drop > 0 && R.r0 > 0 && lr[0] <= R.stop_drop * R.r0) R.reached = true;
  }
  if (steps_done) *steps_done = done;
  return BES_OK;
}

int bes_implicit_run(bes_ctx *c, int nsteps, double cfl_min, double cfl_max, const double *exponents, int local,
                     int repeat_nv, double stop_drop, double *hist, int *steps_done) {
  if (!c || !c->impl || nsteps < 0 || !exponents) { set_error("invalid argument"); return BES_ERR_ARG; }
  Impl *im = impl_of(c);
  Impl::Run &R = im->run;
  R = Impl::Run();
  R.active = true;
  R.cfl = R.cfl_min = cfl_min;
  R.cfl_max = cfl_max;
  R.ex[0] = exponents[0];
  R.ex[1] = exponents[1];
  R.local = local;
  R.repeat_nv = repeat_nv;
  R.stop_drop = stop_drop;
  R.lb_old.assign(c->nv, 0.0);
  return run_steps(c, im, nsteps, hist, steps_done);
}

int bes_implicit_continue(bes_ctx *c, int nsteps, double *hist, int *steps_done, int *reached) {
  if (!c || !c->impl || nsteps < 0 || !impl_of(c)->run.active) {
    set_error("no pseudo-time loop to continue (bes_implicit_run first)");
    return BES_ERR_ARG;
  }
  Impl *im = impl_of(c);
  int s = run_steps(c, im, nsteps, hist, steps_done);
  if (reached) *reached = im->run.reached ? 1 : 0;
  return s;
}

}  // extern "C"
