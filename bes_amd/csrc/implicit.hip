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
#include <chrono>
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

int launch_timestep(bes_ctx *c, const double *U, double cfl, int local, const double *cfl_b = nullptr);
int launch_face_flux(bes_ctx *c);
int launch_gradlim(bes_ctx *c, const double *U, double *Gout, double *Pout);
int launch_residual(bes_ctx *c, const double *U0, double *Ud, double alpha, int local, double *Rout,
                    unsigned long long *linf_bits);
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

// padded length of one Jacobian block record: nv*nv entries rounded up to 16 bytes
template <typename TJ, int NB>
struct JRec {
  static constexpr int P = (NB * (int)sizeof(TJ) + 15) / 16 * 16 / (int)sizeof(TJ);
};

template <int ND, typename TJ>
__global__ __launch_bounds__(256) void k_face_jacobian(const double *__restrict__ U, int N, FaceDev fd,
                                                       GasConst gc, const double *__restrict__ finf,
                                                       TJ *__restrict__ JF, int lgB) {
  constexpr int NV = ND + 2, NB = NV * NV, JP = JRec<TJ, NB>::P;
  // groups of 64 lanes x 2nv input directions: the lanes of a wave take 64/B
  // consecutive faces x B conditions of one direction, and the group's waves
  // reuse the same face states through the cache and fill the same face
  // records; a face's two blocks sit side by side in one record [2][JP] (JL,
  // then JR, row-major), entry e of condition b at (rec + e)*B + b
  const int B = 1 << lgB;
  const int t = xcd_block() * blockDim.x + threadIdx.x;
  const int grp = t / (64 * 2 * NV), rem = t - grp * (64 * 2 * NV);
  const int dir = rem >> 6, l = rem & 63;
  const int f = grp * (64 >> lgB) + (l >> lgB), b = l & (B - 1);
  if (f >= fd.F) return;
  const int o = fd.fo[f], j = fd.fn[f];
  if (j < 0 && dir >= NV) return;   // boundary faces: left Jacobian only
  double nn[ND];
#pragma unroll
  for (int d = 0; d < ND; d++) nn[d] = fd.fnrm[(size_t)d * fd.F + f];
  Dual ul[NV], ur[NV + 1], F[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) ul[v] = Dual(U[((size_t)v * N + o) * B + b], dir == v ? 1.0 : 0.0);
  if (j >= 0) {
#pragma unroll
    for (int v = 0; v < NV; v++) ur[v] = Dual(U[((size_t)v * N + j) * B + b], dir == NV + v ? 1.0 : 0.0);
    roe_flux<ND, Dual>(gc, ul, ur, nn, F);
  } else if (j == -1) {
    slip_flux<ND, Dual>(gc, ul, nn, F);
  } else {
#pragma unroll
    for (int v = 0; v < NV + 1; v++) ur[v] = Dual(finf[b * (NV + 1) + v]);
    freestream_flux<ND, Dual>(gc, ul, ur, nn, F);
  }
  const int side = dir < NV ? 0 : 1, col = dir < NV ? dir : dir - NV;
  TJ *rec = JF + ((size_t)f * 2 + side) * JP * B;
#pragma unroll
  for (int r = 0; r < NV; r++) rec[(size_t)(r * NV + col) * B + b] = (TJ)F[r].d;
}

// one padded block record (16-byte aligned, JP entries) into registers as doubles
template <typename TJ, int NB>
__device__ __forceinline__ void load_rec(const TJ *__restrict__ rec, double *out) {
  constexpr int JP = JRec<TJ, NB>::P, W = 16 / (int)sizeof(TJ);
  TJ tmp[JP];
#pragma unroll
  for (int q = 0; q < JP / W; q++) {
    const uint4 u = reinterpret_cast<const uint4 *>(rec)[q];
    __builtin_memcpy(tmp + q * W, &u, 16);
  }
#pragma unroll
  for (int e = 0; e < NB; e++) out[e] = (double)tmp[e];
}

// ---------------------------------------------------------------------------
// assembly: right-hand side from second-order fluxes, blocks from JF
// ---------------------------------------------------------------------------
template <int ND, typename TM>
__global__ __launch_bounds__(256) void k_assemble(
    const double *__restrict__ FF, int fs, const double *__restrict__ dt, const double *__restrict__ dtg, DevMesh dm,
    FaceDev fd, const TM *__restrict__ JF, int repeat, double *__restrict__ b, TM *__restrict__ A,
    unsigned long long *__restrict__ linf_b, unsigned long long *__restrict__ linf_r, int lgB) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV, JP = JRec<TM, NB>::P;
  const int N = dm.N, B = 1 << lgB;
  const int t = cell_index();
  const int i = t >> lgB, bq = t & (B - 1);
  double bi[NV], ri[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) { bi[v] = 0.0; ri[v] = 0.0; }
  if (i < N) {
    const double mdt = (dtg ? dtg[bq] : dt[(size_t)i * B + bq]) / dm.vol[i];
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
        const double F = FF[((size_t)fdv * fs + v) * B + bq];
        if (own) { bi[v] -= mul * F; ri[v] += F * Af; }
        else { bi[v] += mdt * Af * F; ri[v] -= F * Af; }
      }
    }
#pragma unroll
    for (int v = 0; v < NV; v++) b[((size_t)i * NV + v) * B + bq] = bi[v];
    // blocks: identity, then per face (reference order) the repeated adds.
    // own row: diag += mul*JL, off(i,nbr) = mul*JR; neighbour row:
    // diag += (-m)*JR, off(i,owner) = (-m)*JL.  The diagonal sums in the
    // reference's face order; each off-diagonal block depends on its own face
    // only, so those are written in a second loop in slot order, where every
    // lane of a wave stores into the same slot's planes (whole cache lines;
    // in face order the lanes' slots differ and lines leave L2 part-written).
    double D[NB];
#pragma unroll
    for (int e = 0; e < NB; e++) D[e] = 0.0;
#pragma unroll
    for (int a = 0; a < NV; a++) D[a * NV + a] += 1.0;
    for (int q = 0; q < K; q++) {
      const int m = (info >> (2 * q)) & 3;
      const bool own = (info >> (8 + m)) & 1;
      const int f = fd.cf[(size_t)m * N + i];
      const double Af = dm.af[(size_t)m * N + i];
      const double mul = own ? mdt * Af : -(mdt * Af);
      double J[NB];
      const TM *Jd = JF + ((size_t)f * 2 + (own ? 0 : 1)) * JP * B;
      if (B == 1) load_rec<TM, NB>(Jd, J);
      else {
#pragma unroll
        for (int e = 0; e < NB; e++) J[e] = (double)Jd[(size_t)e * B + bq];
      }
#pragma unroll
      for (int e = 0; e < NB; e++) {
        const double x = mul * J[e];
        for (int r = 0; r < repeat; r++) D[e] += x;
      }
    }
#pragma unroll
    for (int m = 0; m < K; m++) {
      const bool own = (info >> (8 + m)) & 1;
      const int j = dm.nbr[(size_t)m * N + i];
      const int f = fd.cf[(size_t)m * N + i];
      const double Af = dm.af[(size_t)m * N + i];
      const double mul = own ? mdt * Af : -(mdt * Af);
      double J[NB];
      TM *Ao = A + (size_t)(1 + m) * NB * N * B;
      if (j >= 0) {
        const TM *Jo = JF + ((size_t)f * 2 + (own ? 1 : 0)) * JP * B;
        if (B == 1) load_rec<TM, NB>(Jo, J);
        else {
#pragma unroll
          for (int e = 0; e < NB; e++) J[e] = (double)Jo[(size_t)e * B + bq];
        }
#pragma unroll
        for (int e = 0; e < NB; e++) {
          const double x = mul * J[e];
          double s = 0.0;
          for (int r = 0; r < repeat; r++) s += x;
          Ao[((size_t)e * N + i) * B + bq] = (TM)s;
        }
      } else {
#pragma unroll
        for (int e = 0; e < NB; e++) Ao[((size_t)e * N + i) * B + bq] = (TM)0;
      }
    }
#pragma unroll
    for (int e = 0; e < NB; e++) A[((size_t)e * N + i) * B + bq] = (TM)D[e];
  }
  if (linf_b) {   // per condition: [v][B]
#pragma unroll
    for (int v = 0; v < NV; v++) {
      double a = block_max_cond(::fabs(bi[v]), B);
      if ((int)threadIdx.x < B)
        atomicMax(&linf_slot(linf_b, 16 * B)[v * B + threadIdx.x], (unsigned long long)__double_as_longlong(a));
      double c = block_max_cond(::fabs(ri[v]), B);
      if ((int)threadIdx.x < B)
        atomicMax(&linf_slot(linf_r, 16 * B)[v * B + threadIdx.x], (unsigned long long)__double_as_longlong(c));
    }
  }
}

// block Jacobi: LU slot 0 <- inverse of the diagonal block
template <int ND, typename TM>
__global__ __launch_bounds__(256) void k_bj_setup(const TM *__restrict__ A, TM *__restrict__ LU, int N, int lgB) {
  constexpr int NV = ND + 2, NB = NV * NV;
  const int B = 1 << lgB, t = cell_index();
  const int i = t >> lgB, b = t & (B - 1);
  if (i >= N) return;
  double D[NB], Inv[NB];
  load_block_b<TM, NB>(A, 0, N, i, B, b, D);
  block_inverse<NV>(D, Inv);
  store_block_b<TM, NB>(LU, 0, N, i, B, b, Inv);
}

// colour-major copy of the matrix: Ac[se][cpos[i]] = A[se][i], se = slot*nv^2 + e.
// One thread per (cell, condition) copies all nse planes: its reads are
// coalesced, and since the cells of each row group of a colour keep their
// order, the lanes' stores land in a few runs of consecutive positions.  On
// the 2M wing it fetches the matrix once and writes 1.28x it; gathering in
// position order instead read a few entries of each line and fetched 5.7x.
template <typename TM>
__global__ __launch_bounds__(256) void k_ilu_scatter(const TM *__restrict__ A, const int *__restrict__ cpos, int N,
                                                     int nse, TM *__restrict__ Ac, int lgB) {
  const size_t B = (size_t)1 << lgB;
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)N * B) return;
  const size_t i = q >> lgB, b = q & (B - 1), p = (size_t)cpos[i];
#pragma unroll 5
  for (int se = 0; se < nse; se++) Ac[((size_t)se * N + p) * B + b] = A[((size_t)se * N + i) * B + b];
}

// y = A x with matrix and vectors in colour order (neighbour positions cnp), B conditions
template <int ND, typename TM, typename TV>
__global__ __launch_bounds__(256) void k_spmv_pos(const TM *__restrict__ A, const int *__restrict__ cnp, int N,
                                                  const TV *__restrict__ x, TV *__restrict__ y, int lgB) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int B = 1 << lgB;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = t >> lgB, b = t & (B - 1);
  if (p >= N) return;
  // block entries from the row's first entry with the uniform stride es (a
  // full index per entry costs a 64-bit register pair per load in flight)
  const size_t es = (size_t)N * B;
  const TM *Ar = A + (size_t)p * B + b;
  double acc[NV], xv[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) xv[c] = (double)x[((size_t)p * NV + c) * B + b];
#pragma unroll
  for (int r = 0; r < NV; r++) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < NV; c++) s += (double)Ar[(r * NV + c) * es] * xv[c];
    acc[r] = s;
  }
#pragma unroll
  for (int m = 0; m < K; m++) {
    const int q = cnp[(size_t)m * N + p];
    if (q < 0) continue;
#pragma unroll
    for (int c = 0; c < NV; c++) xv[c] = (double)x[((size_t)q * NV + c) * B + b];
    const TM *Ao = Ar + (size_t)(1 + m) * NB * es;
#pragma unroll
    for (int r = 0; r < NV; r++) {
      double s = 0;
#pragma unroll
      for (int c = 0; c < NV; c++) s += (double)Ao[(r * NV + c) * es] * xv[c];
      acc[r] += s;
    }
  }
#pragma unroll
  for (int r = 0; r < NV; r++) y[((size_t)p * NV + r) * B + b] = (TV)acc[r];
}

// rows between cell order and colour order (each row nv*B contiguous values):
// to_colour: dst[p] = src[crow[p]], else dst[crow[p]] = src[p]
template <typename T>
__global__ void k_perm_rows(const T *__restrict__ src, const int *__restrict__ crow, int N, int rw, int to_colour,
                            T *__restrict__ dst) {
  const size_t tot = (size_t)N * rw;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x) {
    const size_t p = q / rw, c = q - p * rw;
    const size_t i = (size_t)crow[p];
    if (to_colour) dst[q] = src[i * rw + c];
    else dst[i * rw + c] = src[q];
  }
}

// block ILU(0) of the rows at positions p0 .. p0+nrows-1 (one colour); rows of
// earlier colours are final.  Elimination order: colour, then cell index (not
// position, so the factors do not depend on how rows are laid out in a colour).
// A row's thread first copies its blocks from the colour-order matrix Ac into
// LU (only that thread ever touches the row), eliminates, and leaves slot 0 =
// inv(U_ii), lower slots L_ik, upper slots U_ij.  Finished rows are also kept
// row-major by position (Dr: inv(U_pp); Ur: upper blocks by slot), so a later
// row reads a neighbour's block as one contiguous record.
template <int ND, typename TM>
__global__ __launch_bounds__(256) void k_ilu_factor_color(const TM *__restrict__ Ac, TM *__restrict__ LU, int N,
                                                          const int *__restrict__ cnb, const int *__restrict__ cnp,
                                                          const unsigned char *__restrict__ clow,
                                                          const int *__restrict__ color, const int *__restrict__ crow,
                                                          TM *__restrict__ Dr, TM *__restrict__ Ur, int p0, int nrows,
                                                          int lgB) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int B = 1 << lgB;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (nrows << lgB)) return;
  const int p = p0 + (t >> lgB), b = t & (B - 1);
  const int i = crow[p];
  const unsigned lowm = clow[p];
  int nb[K], np_[K];
#pragma unroll
  for (int m = 0; m < K; m++) {
    nb[m] = cnb[(size_t)m * N + p];
    np_[m] = cnp[(size_t)m * N + p];
  }
  // the row's own blocks into the factor storage (boundary slots stay zero),
  // a block at a time (unrolled over the slots, the compiler would hold all
  // 125 entries in registers between the loads and the stores)
#pragma unroll 1
  for (int sl = 0; sl <= K; sl++) {
    if (sl > 0 && nb[sl - 1] < 0) continue;
#pragma unroll
    for (int e = 0; e < NB; e++) {
      const size_t q = (((size_t)sl * NB + e) * N + p) * B + b;
      LU[q] = Ac[q];
    }
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
  load_block_b<TM, NB>(LU, 0, N, p, B, b, Aii);
  const size_t es = (size_t)N * B;   // stride between a block's entries
  // Each product below sums over pp in the same order as a whole-block
  // product would; they are arranged so that few values are live at once (one
  // row of A_ik, one column of U_kj), which keeps the kernel at two waves per
  // SIMD instead of one.
  for (int q = 0; q < nl; q++) {
    const int m = ord[q];
    const int pk = np_[m];
    double Lik[NB];
    {
      // row k is final: inv(U_kk), from the row-major copy (one contiguous block)
      double Uinv[NB];
#pragma unroll
      for (int e = 0; e < NB; e++) Uinv[e] = (double)Dr[((size_t)pk * NB + e) * B + b];
      // (entries addressed from a per-row base with a uniform stride: computing
      // each entry's full index costs a 64-bit register pair per load in flight)
      const TM *aik = LU + ((size_t)(1 + m) * NB * N + p) * B + b;
#pragma unroll
      for (int r = 0; r < NV; r++) {
        double ar[NV];   // row r of A_ik
#pragma unroll
        for (int pp = 0; pp < NV; pp++) ar[pp] = (double)aik[(r * NV + pp) * es];
#pragma unroll
        for (int c = 0; c < NV; c++) {
          double s = 0;
#pragma unroll
          for (int pp = 0; pp < NV; pp++) s += ar[pp] * Uinv[pp * NV + c];
          Lik[r * NV + c] = s;
        }
      }
    }
    store_block_b<TM, NB>(LU, 1 + m, N, p, B, b, Lik);
    // A_ij -= L_ik U_kj for j in row k's upper part that are in row i's pattern
    // (one j at a time: unrolled, the U_kj loads of all j would be in flight)
    const unsigned lowk = clow[pk];
#pragma unroll 1
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
      const TM *urec = Ur + ((size_t)pk * K + mk) * NB * B + b;
#pragma unroll
      for (int c = 0; c < NV; c++) {
        double uc[NV];   // column c of U_kj
#pragma unroll
        for (int pp = 0; pp < NV; pp++) uc[pp] = (double)urec[(size_t)(pp * NV + c) * B];
#pragma unroll
        for (int r = 0; r < NV; r++) {
          double s = 0;
#pragma unroll
          for (int pp = 0; pp < NV; pp++) s += Lik[r * NV + pp] * uc[pp];
          if (target == -2) {
            Aii[r * NV + c] -= s;
          } else {
            TM *aij = LU + ((size_t)(1 + target) * NB * N + p) * B + b + (r * NV + c) * es;
            *aij = (TM)((double)*aij - s);
          }
        }
      }
    }
  }
  double Inv[NB];
  block_inverse<NV>(Aii, Inv);
  store_block_b<TM, NB>(LU, 0, N, p, B, b, Inv);
#pragma unroll
  for (int e = 0; e < NB; e++) Dr[((size_t)p * NB + e) * B + b] = (TM)Inv[e];
  // upper blocks, final from here on, row-major for the rows of later colours
  // (a block at a time, as the copy above)
#pragma unroll 1
  for (int m = 0; m < K; m++) {
    if (nb[m] < 0 || ((lowm >> m) & 1)) continue;
#pragma unroll
    for (int e = 0; e < NB; e++)
      Ur[(((size_t)p * K + m) * NB + e) * B + b] = LU[(((size_t)(1 + m) * NB + e) * N + p) * B + b];
  }
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
                                                   TV *__restrict__ x, int lgB) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int B = 1 << lgB;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (nrows << lgB)) return;
  const int p = p0 + (t >> lgB), b = t & (B - 1);
  const unsigned lowm = clow[p];
  const size_t es = (size_t)N * B;   // stride between a block's entries
  const TM *Lr = LU + (size_t)p * B + b;
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) acc[c] = (double)r[((size_t)p * NV + c) * B + b];
#pragma unroll
  for (int m = 0; m < K; m++) {
    if (!((lowm >> m) & 1)) continue;
    const int q = cnp[(size_t)m * N + p];
    const TV *src = q < y0end ? r : y;
    double yk[NV];
#pragma unroll
    for (int c = 0; c < NV; c++) yk[c] = (double)src[((size_t)q * NV + c) * B + b];
    const TM *L = Lr + (size_t)(1 + m) * NB * es;
#pragma unroll
    for (int a = 0; a < NV; a++) {
      double s = 0;
#pragma unroll
      for (int c = 0; c < NV; c++) s += (double)L[(a * NV + c) * es] * yk[c];
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
      for (int c = 0; c < NV; c++) s += (double)Lr[(a * NV + c) * es] * acc[c];
      x[((size_t)p * NV + a) * B + b] = (TV)s;
    }
  } else {
#pragma unroll
    for (int c = 0; c < NV; c++) y[((size_t)p * NV + c) * B + b] = (TV)acc[c];
  }
}

// backward sweep of one colour (vectors in colour order): x_p = inv(U_pp) (y_p - sum_(j upper) U_pj x_j);
// y_p is r_p below y0end (colour 0)
template <int ND, typename TM, typename TV>
__global__ __launch_bounds__(256) void k_ilu_upper(const TM *__restrict__ LU, int N, const int *__restrict__ cnp,
                                                   const unsigned char *__restrict__ clow, int p0, int nrows,
                                                   const TV *__restrict__ y, TV *__restrict__ x,
                                                   const TV *__restrict__ r, int y0end, int lgB) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int B = 1 << lgB;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (nrows << lgB)) return;
  const int p = p0 + (t >> lgB), b = t & (B - 1);
  const unsigned lowm = clow[p];
  const TV *own = p < y0end ? r : y;
  const size_t es = (size_t)N * B;   // stride between a block's entries
  const TM *Lr = LU + (size_t)p * B + b;
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) acc[c] = (double)own[((size_t)p * NV + c) * B + b];
#pragma unroll
  for (int m = 0; m < K; m++) {
    const int q = cnp[(size_t)m * N + p];
    if (q < 0 || ((lowm >> m) & 1)) continue;
    double xj[NV];
#pragma unroll
    for (int c = 0; c < NV; c++) xj[c] = (double)x[((size_t)q * NV + c) * B + b];
    const TM *Uo = Lr + (size_t)(1 + m) * NB * es;
#pragma unroll
    for (int a = 0; a < NV; a++) {
      double s = 0;
#pragma unroll
      for (int c = 0; c < NV; c++) s += (double)Uo[(a * NV + c) * es] * xj[c];
      acc[a] -= s;
    }
  }
#pragma unroll
  for (int a = 0; a < NV; a++) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < NV; c++) s += (double)Lr[(a * NV + c) * es] * acc[c];
    x[((size_t)p * NV + a) * B + b] = (TV)s;
  }
}

// ---------------------------------------------------------------------------
// Krylov kernels (vectors [N][nv][B], type TV)
// ---------------------------------------------------------------------------
// z = inv(D_i) r_i (block Jacobi)
template <int ND, typename TM, typename TV>
__global__ __launch_bounds__(256) void k_bj_apply(const TM *__restrict__ LU, int N, const TV *__restrict__ r,
                                                  TV *__restrict__ z, int lgB) {
  constexpr int NV = ND + 2;
  const int B = 1 << lgB, t = cell_index();
  const int i = t >> lgB, b = t & (B - 1);
  if (i >= N) return;
  double ri[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) ri[c] = (double)r[((size_t)i * NV + c) * B + b];
#pragma unroll
  for (int a = 0; a < NV; a++) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < NV; c++) s += (double)LU[((size_t)(a * NV + c) * N + i) * B + b] * ri[c];
    z[((size_t)i * NV + a) * B + b] = (TV)s;
  }
}

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

// Deterministic reductions over a batch of B conditions (vectors [n1][B],
// condition fastest).  Every condition's n1 elements are summed in one fixed
// tree whatever B is, so a condition's dot products, norms and hence its whole
// Krylov solve are the same bits in a batch as alone: lane l of L (L depends
// on n1 only) adds, in order, the elements of the groups g = l, l + L, l + 2L,
// ..., group g being elements 4g .. 4g + 3 (zeros past n1); the L lane sums of
// a condition are then added by sum_lanes in a fixed two-level tree.  Thread
// l*B + b is lane l of condition b, so the loads of a wave are contiguous for
// any B, and one 16-byte packet (FP32) or two (FP64) per vector when B = 1.
constexpr int kLaneGroup = 4;

template <typename TV>
__device__ __forceinline__ void ld_group(const TV *__restrict__ v, size_t e, size_t n1, int lgB, int b,
                                         double x[kLaneGroup]) {
  if (lgB == 0 && e + kLaneGroup <= n1) {
    if (sizeof(TV) == 8) {
      const double2 a = *reinterpret_cast<const double2 *>(v + e);
      const double2 c = *reinterpret_cast<const double2 *>(v + e + 2);
      x[0] = a.x; x[1] = a.y; x[2] = c.x; x[3] = c.y;
    } else {
      const float4 a = *reinterpret_cast<const float4 *>(v + e);
      x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
    }
  } else {
#pragma unroll
    for (int h = 0; h < kLaneGroup; h++) x[h] = e + h < n1 ? (double)v[((e + h) << lgB) + b] : 0.0;
  }
}

template <typename TV>
__device__ __forceinline__ void st_group(TV *__restrict__ v, size_t e, size_t n1, int lgB, int b,
                                         const TV r[kLaneGroup]) {
  if (lgB == 0 && e + kLaneGroup <= n1) {
    if (sizeof(TV) == 8) {
      *reinterpret_cast<double2 *>(v + e) = make_double2((double)r[0], (double)r[1]);
      *reinterpret_cast<double2 *>(v + e + 2) = make_double2((double)r[2], (double)r[3]);
    } else {
      *reinterpret_cast<float4 *>(v + e) = make_float4((float)r[0], (float)r[1], (float)r[2], (float)r[3]);
    }
  } else {
#pragma unroll
    for (int h = 0; h < kLaneGroup; h++)
      if (e + h < n1) v[((e + h) << lgB) + b] = r[h];
  }
}

// up to 8 dot products <V_(i0+q), w> and ||w||^2 per lane: S[(q*B + b)*L + l],
// the norm in slot cnt
template <typename TV>
__global__ __launch_bounds__(256) void k_mdot(const TV *__restrict__ V, size_t stride, int i0, int cnt,
                                              const TV *__restrict__ w, size_t n1, int lgB, int L,
                                              double *__restrict__ S) {
  const int B = 1 << lgB;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ((size_t)L << lgB)) return;
  const size_t l = g >> lgB;
  const int b = (int)(g & (B - 1));
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; q++) acc[q] = 0;
  for (size_t e = kLaneGroup * l; e < n1; e += (size_t)kLaneGroup * L) {
    double wx[kLaneGroup];
    ld_group(w, e, n1, lgB, b, wx);
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (q < cnt) {
        double vx[kLaneGroup];
        ld_group(V + (size_t)(i0 + q) * stride, e, n1, lgB, b, vx);
#pragma unroll
        for (int h = 0; h < kLaneGroup; h++) acc[q] += vx[h] * wx[h];
      }
#pragma unroll
    for (int h = 0; h < kLaneGroup; h++) acc[8] += wx[h] * wx[h];
  }
  for (int q = 0; q < cnt; q++) S[((size_t)q * B + b) * L + l] = acc[q];
  S[((size_t)cnt * B + b) * L + l] = acc[8];
}

// level 1 of the lane sums: block (x, qb) adds lanes x*1024 + t + 256k of row
// qb (k = 0..3 in order, thread t), then a fixed tree over the block.  With
// one block per row (L <= 1024) the result is final and goes to
// out[q'*out_stride_q + b] (qb = q*B + b; q' = 8 for q = qnorm, else q);
// otherwise to S2[qb*G + x] for level 2.
__global__ __launch_bounds__(256) void k_sum_lanes1(const double *__restrict__ S, int L, double *__restrict__ dst,
                                                    int out_stride_q, int qnorm, int lgB) {
  const int x = blockIdx.x, qb = blockIdx.y, G = gridDim.x, B = 1 << lgB;
  double s = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int l = x * 1024 + k * 256 + (int)threadIdx.x;
    if (l < L) s += S[(size_t)qb * L + l];
  }
  __shared__ double sh[4];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double r = sh[0] + sh[1] + sh[2] + sh[3];
  if (G > 1) {
    dst[(size_t)qb * G + x] = r;
  } else {
    const int q = qb >> lgB, qo = q == qnorm ? 8 : q;
    dst[(size_t)qo * out_stride_q + (qb & (B - 1))] = r;
  }
}

// level 2: out row qb = sum of its G level-1 sums, thread t adding t, t + 256,
// ... in order, then a fixed tree (same output mapping as level 1)
__global__ __launch_bounds__(256) void k_sum_lanes(const double *__restrict__ S, int L, double *__restrict__ out,
                                                   int out_stride_q, int qnorm, int lgB) {
  const int qb = blockIdx.x, B = 1 << lgB;
  double s = 0;
  for (int l = threadIdx.x; l < L; l += 256) s += S[(size_t)qb * L + l];
  __shared__ double sh[4];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int q = qb >> lgB, qo = q == qnorm ? 8 : q;
    out[(size_t)qo * out_stride_q + (qb & (B - 1))] = sh[0] + sh[1] + sh[2] + sh[3];
  }
}

// w -= sum_q h_q V_q per condition (h_q of condition b at h[(9*(q/8) + q%8)*B + b],
// the multidot's layout) and ||w||^2 per lane into S[b*L + l]
template <typename TV>
__global__ __launch_bounds__(256) void k_orth_update(const TV *__restrict__ V, size_t stride, int cnt,
                                                     const double *__restrict__ h, TV *__restrict__ w, size_t n1,
                                                     int lgB, int L, double *__restrict__ S) {
  const int B = 1 << lgB;
  __shared__ double hs[32 * 64];
  for (int q = threadIdx.x; q < cnt * B; q += blockDim.x) {
    const int qq = q >> lgB, b = q & (B - 1);
    hs[q] = h[(size_t)(9 * (qq >> 3) + (qq & 7)) * B + b];
  }
  __syncthreads();
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ((size_t)L << lgB)) return;
  const size_t l = g >> lgB;
  const int b = (int)(g & (B - 1));
  double nrm = 0;
  for (size_t e = kLaneGroup * l; e < n1; e += (size_t)kLaneGroup * L) {
    double wx[kLaneGroup];
    ld_group(w, e, n1, lgB, b, wx);
    int q = 0;
    for (; q + 4 <= cnt; q += 4) {   // four vectors' loads in flight at once
      double vx[4][kLaneGroup];
#pragma unroll
      for (int u = 0; u < 4; u++) ld_group(V + (size_t)(q + u) * stride, e, n1, lgB, b, vx[u]);
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const double hq = hs[((q + u) << lgB) + b];
#pragma unroll
        for (int k = 0; k < kLaneGroup; k++) wx[k] -= hq * vx[u][k];
      }
    }
    for (; q < cnt; q++) {
      double vx[kLaneGroup];
      ld_group(V + (size_t)q * stride, e, n1, lgB, b, vx);
      const double hq = hs[(q << lgB) + b];
#pragma unroll
      for (int k = 0; k < kLaneGroup; k++) wx[k] -= hq * vx[k];
    }
    TV r[kLaneGroup];
#pragma unroll
    for (int k = 0; k < kLaneGroup; k++) r[k] = (TV)wx[k];
    st_group(w, e, n1, lgB, b, r);
#pragma unroll
    for (int k = 0; k < kLaneGroup; k++) nrm += (double)r[k] * (double)r[k];
  }
  S[(size_t)b * L + l] = nrm;
}

// r = b - y (b FP64) and ||r||^2 per lane into S[b*L + l]
template <typename TV>
__global__ __launch_bounds__(256) void k_resid(const double *__restrict__ bv, const TV *__restrict__ y,
                                               TV *__restrict__ r, size_t n1, int lgB, int L, double *__restrict__ S) {
  const int B = 1 << lgB;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ((size_t)L << lgB)) return;
  const size_t l = g >> lgB;
  const int b = (int)(g & (B - 1));
  double nrm = 0;
  for (size_t e = kLaneGroup * l; e < n1; e += (size_t)kLaneGroup * L) {
    double bx[kLaneGroup], yx[kLaneGroup];
    ld_group(bv, e, n1, lgB, b, bx);
    ld_group(y, e, n1, lgB, b, yx);
    TV rr[kLaneGroup];
#pragma unroll
    for (int k = 0; k < kLaneGroup; k++) {
      const double v = bx[k] - yx[k];
      rr[k] = (TV)v;
      nrm += v * v;
    }
    st_group(r, e, n1, lgB, b, rr);
  }
  S[(size_t)b * L + l] = nrm;
}

// Element-wise vector kernels.  A thread works on groups of 4 consecutive
// elements (one 16-byte packet in FP32, two in FP64; every vector starts on a
// 64-element boundary) in a grid-stride loop; the grid's thread count is a
// multiple of 64, so element t of every group a thread touches belongs to the
// same condition, (4*tid + t) mod B, and per-condition factors are read once.
template <typename T>
__device__ __forceinline__ void ld4(const T *p, double x[4]) {
  if (sizeof(T) == 8) {
    const double2 a = *reinterpret_cast<const double2 *>(p);
    const double2 c = *reinterpret_cast<const double2 *>(p + 2);
    x[0] = a.x; x[1] = a.y; x[2] = c.x; x[3] = c.y;
  } else {
    const float4 a = *reinterpret_cast<const float4 *>(p);
    x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
  }
}
template <typename T>
__device__ __forceinline__ void st4(T *p, const double x[4]) {
  if (sizeof(T) == 8) {
    *reinterpret_cast<double2 *>(p) = make_double2(x[0], x[1]);
    *reinterpret_cast<double2 *>(p + 2) = make_double2(x[2], x[3]);
  } else {
    *reinterpret_cast<float4 *>(p) = make_float4((float)x[0], (float)x[1], (float)x[2], (float)x[3]);
  }
}

// w = w / sqrt(n2[b]) per condition when n2[b] > 0 (the norm read on the
// device, so the launch need not wait for the host)
template <typename TV>
__global__ void k_scale_norm(TV *__restrict__ w, const double *__restrict__ n2, size_t n, int lgB) {
  const size_t B = (size_t)1 << lgB, tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x,
               nt = (size_t)gridDim.x * blockDim.x, ng = n / 4;
  double a[4];
  bool on[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const double s2 = n2[(4 * tid + t) & (B - 1)];
    on[t] = s2 > 0;
    a[t] = on[t] ? 1.0 / sqrt(s2) : 1.0;
  }
  for (size_t g = tid; g < ng; g += nt) {
    double x[4];
    ld4(w + 4 * g, x);
#pragma unroll
    for (int t = 0; t < 4; t++) x[t] = on[t] ? (double)(TV)(a[t] * x[t]) : x[t];
    st4(w + 4 * g, x);
  }
  for (size_t e = 4 * ng + tid; e < n; e += nt) {
    const double s2 = n2[e & (B - 1)];
    if (s2 > 0) w[e] = (TV)((1.0 / sqrt(s2)) * (double)w[e]);
  }
}

// y = a_b x (scale/copy with conversion), a per condition; y may be x
template <typename TA, typename TB>
__global__ void k_scale_b(const TA *x, CoefB a, TB *y, size_t n, int lgB) {
  const size_t B = (size_t)1 << lgB, tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x,
               nt = (size_t)gridDim.x * blockDim.x, ng = n / 4;
  double c[4];
#pragma unroll
  for (int t = 0; t < 4; t++) c[t] = a.c[(4 * tid + t) & (B - 1)];
  for (size_t g = tid; g < ng; g += nt) {
    double v[4];
    ld4(x + 4 * g, v);
#pragma unroll
    for (int t = 0; t < 4; t++) v[t] = c[t] * v[t];
    st4(y + 4 * g, v);
  }
  for (size_t e = 4 * ng + tid; e < n; e += nt) y[e] = (TB)(a.c[e & (B - 1)] * (double)x[e]);
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

// y = a*x (scale/copy with conversion)
template <typename TA, typename TB>
__global__ void k_scale(const TA *__restrict__ x, double a, TB *__restrict__ y, size_t n) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x,
               ng = n / 4;
  for (size_t g = tid; g < ng; g += nt) {
    double v[4];
    ld4(x + 4 * g, v);
#pragma unroll
    for (int t = 0; t < 4; t++) v[t] = a * v[t];
    st4(y + 4 * g, v);
  }
  for (size_t e = 4 * ng + tid; e < n; e += nt) y[e] = (TB)(a * (double)x[e]);
}

// u = sum_q c[q][b] V[q] per condition
template <typename TV>
__global__ void k_combine(const TV *__restrict__ V, size_t stride, int cnt, const double *__restrict__ c,
                          TV *__restrict__ u, size_t n, int lgB) {
  const size_t B = (size_t)1 << lgB, tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x,
               nt = (size_t)gridDim.x * blockDim.x, ng = n / 4;
  int bt[4];
#pragma unroll
  for (int t = 0; t < 4; t++) bt[t] = (int)((4 * tid + t) & (B - 1));
  for (size_t g = tid; g < ng; g += nt) {
    double s[4] = {0, 0, 0, 0};
    for (int q = 0; q < cnt; q++) {
      double v[4];
      ld4(V + (size_t)q * stride + 4 * g, v);
#pragma unroll
      for (int t = 0; t < 4; t++) s[t] += c[q * B + bt[t]] * v[t];
    }
    st4(u + 4 * g, s);
  }
  for (size_t e = 4 * ng + tid; e < n; e += nt) {
    const size_t b = e & (B - 1);
    double s = 0;
    for (int q = 0; q < cnt; q++) s += c[q * B + b] * (double)V[(size_t)q * stride + e];
    u[e] = (TV)s;
  }
}

// x (FP64) += z
template <typename TV>
__global__ void k_axpy_x(double *__restrict__ x, const TV *__restrict__ z, size_t n) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x,
               ng = n / 4;
  for (size_t g = tid; g < ng; g += nt) {
    double xv[4], zv[4];
    ld4(x + 4 * g, xv);
    ld4(z + 4 * g, zv);
#pragma unroll
    for (int t = 0; t < 4; t++) xv[t] += zv[t];
    st4(x + 4 * g, xv);
  }
  for (size_t e = 4 * ng + tid; e < n; e += nt) x[e] += (double)z[e];
}

// implicit update U[v][i][b] += x[i][v][b] for the conditions still advancing
template <int ND>
__global__ __launch_bounds__(256) void k_update(double *__restrict__ U, const double *__restrict__ x, int N, int lgB,
                                                CoefB active) {
  constexpr int NV = ND + 2;
  const int B = 1 << lgB, t = cell_index();
  const int i = t >> lgB, b = t & (B - 1);
  if (i >= N || active.c[b] == 0.0) return;
#pragma unroll
  for (int v = 0; v < NV; v++) U[((size_t)v * N + i) * B + b] += x[((size_t)i * NV + v) * B + b];
}

// bits[q] = max over the Linf copies (copy 0 in place)
__global__ void k_linf_reduce_i(unsigned long long *bits, int nq) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nq) {
    const unsigned long long m = linf_slots_max(bits, nq, q);
    bits[q] = m;
  }
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

// Multicolour order of the block ILU(0) (rows of one colour are never
// coupled, so each colour is one parallel sweep).  The colours come from a
// greedy pass over the cells in an order set by mode: 0 device (Morton) order,
// 1 the mesh file's cell order, 2 along the freestream velocity (increasing
// centroid . u_inf).  The order changes the factors (the elimination order is
// colour, then cell index) and so the Krylov iteration counts, not the method.
static int build_colouring(bes_ctx *c, Impl *im, int mode) {
  const Geometry &g = c->geo;
  const int N = c->N, k = c->k, nd = c->nd;
  std::vector<int> visit(N);
  for (int q = 0; q < N; q++) visit[q] = mode == 1 ? (int)c->iperm[q] : q;
  if (mode == 2) {
    std::vector<double> key(N);
    for (int i = 0; i < N; i++) {
      const int64_t e = c->perm[i];
      double t = 0;
      for (int d = 0; d < nd; d++) t += g.centroid[e * nd + d] * c->finf[1 + d];
      key[i] = t;
    }
    std::stable_sort(visit.begin(), visit.end(), [&](int a, int b) { return key[a] < key[b]; });
  }
  std::vector<int> color(N, -1);
  int ncol = 0;
  for (int q = 0; q < N; q++) {
    const int i = visit[q];
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
  auto up = [&](auto **p, const auto &h) -> int {
    using T = typename std::decay<decltype(h)>::type::value_type;
    if (*p) BES_HIP(hipFree(*p));
    *p = nullptr;
    BES_HIP(hipMalloc((void **)p, std::max<size_t>(h.size(), 1) * sizeof(T)));
    BES_HIP(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return BES_OK;
  };
  int s;
  if ((s = up(&im->color, color)) || (s = up(&im->crow, crow)) || (s = up(&im->cpos, cpos)) ||
      (s = up(&im->cnb, cnb)) || (s = up(&im->cnp, cnp)) || (s = up(&im->clow, clow)))
    return s;
  im->coloring = mode;
  im->factored = false;
  return BES_OK;
}

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
      (s = up(&im->fd.cf, cf)) || (s = up(&im->fd.mrev, mrev)) || (s = build_colouring(c, im, im->coloring)))
    return s;
  const size_t nvec = (size_t)N * nv;
  const size_t es = im->prec == 32 ? sizeof(float) : sizeof(double);
  const size_t mat = (size_t)(k + 1) * NB * N;
  // face Jacobian records [F][2][JP] in the matrix precision (FP32 in the mixed forms)
  const size_t jrec = im->prec == 32 ? (size_t)(nd == 2 ? JRec<float, 16>::P : JRec<float, 25>::P)
                                     : (size_t)(nd == 2 ? JRec<double, 16>::P : JRec<double, 25>::P);
  const size_t Bs = (size_t)c->B;   // conditions of the batch, the fastest index of every array below
  BES_HIP(hipMalloc((void **)&im->JF, (size_t)F * 2 * jrec * es * Bs));
  BES_HIP(hipMemset(im->JF, 0, (size_t)F * 2 * jrec * es * Bs));
  BES_HIP(hipMalloc((void **)&im->b, nvec * Bs * sizeof(double)));
  BES_HIP(hipMalloc((void **)&im->x, nvec * Bs * sizeof(double)));
  BES_HIP(hipMalloc(&im->A, mat * es * Bs));
  BES_HIP(hipMalloc(&im->LU, mat * es * Bs));
  BES_HIP(hipMalloc(&im->Ac, mat * es * Bs));
  BES_HIP(hipMalloc((void **)&im->bc, nvec * Bs * sizeof(double)));
  BES_HIP(hipMalloc((void **)&im->xc, nvec * Bs * sizeof(double)));
  im->vstride = (nvec * Bs + 63) / 64 * 64;   // 16-byte aligned packet loads on every vector
  // lanes of the deterministic reductions: a power of two set by the size of
  // one condition's vector only (each lane adds 16 to 64 elements; 2^18 lanes,
  // a thread each, for the 2M-cell wing)
  im->lanes = 256;
  while (im->lanes < (1 << 20) && (size_t)im->lanes * 64 < nvec) im->lanes *= 2;
  // GMRES(m) keeps m + 1 basis vectors; BiCGStab's seven vectors use the same storage
  BES_HIP(hipMalloc(&im->V, (size_t)std::max(im->restart + 1, 7) * im->vstride * es));
  BES_HIP(hipMalloc(&im->Z, 3 * im->vstride * es));
  BES_HIP(hipMalloc((void **)&im->part, 32 * kRedBlocks * sizeof(double)));
  BES_HIP(hipMalloc((void **)&im->S, (size_t)9 * Bs * im->lanes * sizeof(double)));
  BES_HIP(hipMalloc((void **)&im->S2, (size_t)9 * Bs * ((im->lanes + 1023) / 1024) * sizeof(double)));
  BES_HIP(hipMalloc((void **)&im->dots, 64 * Bs * sizeof(double)));
  BES_HIP(hipHostMalloc((void **)&im->hdots, 64 * Bs * sizeof(double), hipHostMallocDefault));
  BES_HIP(hipEventCreateWithFlags(&im->ev_dots, hipEventDisableTiming));
  BES_HIP(hipMalloc((void **)&im->bits, (size_t)kLinfSlots * 16 * Bs * sizeof(unsigned long long)));
  return BES_OK;
}

static void impl_release(Impl *im) {
  lp::lp_release(im);
  amg_release(im);
  void *ptrs[] = {im->fd.fo, im->fd.fn, im->fd.fnrm, im->fd.cf, im->fd.mrev, im->color, im->crow, im->JF, im->S, im->S2,
                  im->b, im->x, im->A, im->LU, im->LM, im->LMh, im->V, im->Z, im->part, im->dots, im->bits, im->Ac,
                  im->bc, im->xc,
                  im->cpos, im->cnb, im->cnp, im->clow, im->Dr, im->Ur};
  for (void *p : ptrs)
    if (p) hipFree(p);
  if (im->hdots) hipHostFree(im->hdots);
  if (im->ev_dots) hipEventDestroy(im->ev_dots);
  delete im;
}

template <typename TM>
static int launch_assemble(bes_ctx *c, Impl *im) {
  const int bs = tuned_bs(c, kProfAssemble), bsj = tuned_bs(c, kProfJacobian), B = c->B, lgB = c->lgB,
            nb = grid_for(c->N * B, bs);
  const double *dtg = c->dt_local ? nullptr : c->dtg;
  const double nbl = (double)c->nv * c->nv;
  {
    // per face: both states, normal, ids; writes one or two nv x nv blocks
    ProfScope ps(c, kProfJacobian, (double)im->fd.F * (16.0 * c->nv * B + 8.0 * c->nd + 8.0) +
                                       (double)(im->fd.F + im->ninternal) * nbl * sizeof(TM) * B);
    const int groups = (im->fd.F + (64 >> lgB) - 1) / (64 >> lgB);   // groups of 64 lanes
    const int nt = groups * 64 * 2 * c->nv;
    if (c->nd == 2)
      hipLaunchKernelGGL((k_face_jacobian<2, TM>), dim3(grid_for(nt, bsj)), dim3(bsj), 0, c->stream, c->U0, c->N,
                         im->fd, c->gc, c->dfinf, (TM *)im->JF, lgB);
    else
      hipLaunchKernelGGL((k_face_jacobian<3, TM>), dim3(grid_for(nt, bsj)), dim3(bsj), 0, c->stream, c->U0, c->N,
                         im->fd, c->gc, c->dfinf, (TM *)im->JF, lgB);
    BES_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_fill_u64_i, dim3((kLinfSlots * 16 * B + 63) / 64), dim3(64), 0, c->stream, im->bits,
                     kLinfSlots * 16 * B, 0ull);
  int s = launch_face_flux(c);
  if (s) return s;
  // per cell: dt, volume, face records, the cell's face fluxes and both
  // Jacobian blocks of each face; writes b and k+1 matrix blocks
  // (dt, volume, face order word; per face: face id twice, area, neighbour;
  // b written), the face fluxes, every face Jacobian block once, the k+1
  // blocks of every row written
  ProfScope ps(c, kProfAssemble,
               (double)c->N * (12.0 + 8.0 * B + c->k * 20.0 + 8.0 * c->nv * B) +
                   (double)c->dm.F * 8.0 * c->fs * B + (double)(im->fd.F + im->ninternal) * nbl * sizeof(TM) * B +
                   (double)c->N * (c->k + 1) * nbl * sizeof(TM) * B);
  if (c->nd == 2)
    hipLaunchKernelGGL((k_assemble<2, TM>), dim3(nb), dim3(bs), 0, c->stream, c->FF, c->fs, c->dt, dtg, c->dm, im->fd,
                       (const TM *)im->JF, im->repeat_nv, im->b, (TM *)im->A, im->bits, im->bits + 8 * B, lgB);
  else
    hipLaunchKernelGGL((k_assemble<3, TM>), dim3(nb), dim3(bs), 0, c->stream, c->FF, c->fs, c->dt, dtg, c->dm, im->fd,
                       (const TM *)im->JF, im->repeat_nv, im->b, (TM *)im->A, im->bits, im->bits + 8 * B, lgB);
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
  const int bs = tuned_bs(c, kProfFactor), N = c->N, nv = c->nv, B = c->B, lgB = c->lgB;
  const double used = (double)(N + 2 * im->ninternal) * nv * nv * sizeof(TM) * B;   // stored blocks
  // ILU(0): the matrix read and its colour-order copy written, that copy read
  // and the factors written, the row-major pivot inverses and upper blocks
  // written and read once each
  const double rowmaj = (double)(N + im->ninternal) * nv * nv * sizeof(TM) * B;
  ProfScope ps(c, kProfFactor, im->precond == 1 ? 2.0 * N * nv * nv * sizeof(TM) * B : 4.0 * used + 2.0 * rowmaj);
  if (im->precond == 3) {
    const int s = lp::lp_build(c, im);
    if (s) return s;
    im->lp_since = 0;
    im->lp_current = true;
    im->lp_builds++;
    return BES_OK;
  }
  if (im->precond == 1) {
    if (c->nd == 2) hipLaunchKernelGGL((k_bj_setup<2, TM>), dim3(grid_for(N * B, bs)), dim3(bs), 0, c->stream, (const TM *)im->A, (TM *)im->LU, N, lgB);
    else hipLaunchKernelGGL((k_bj_setup<3, TM>), dim3(grid_for(N * B, bs)), dim3(bs), 0, c->stream, (const TM *)im->A, (TM *)im->LU, N, lgB);
  } else if (im->precond == 2) {
    if (!im->Dr) BES_HIP(hipMalloc(&im->Dr, (size_t)N * nv * nv * B * sizeof(TM)));
    if (!im->Ur) BES_HIP(hipMalloc(&im->Ur, (size_t)N * c->k * nv * nv * B * sizeof(TM)));
    hipLaunchKernelGGL((k_ilu_scatter<TM>), dim3(grid_for(N * B, 256)), dim3(256), 0, c->stream, (const TM *)im->A,
                       im->cpos, N, (c->k + 1) * nv * nv, (TM *)im->Ac, lgB);
    for (int col = 0; col < im->ncolor; col++) {
      const int r0 = im->color_start[col], nr = im->color_start[col + 1] - r0;
      if (nr == 0) continue;
      if (c->nd == 2)
        hipLaunchKernelGGL((k_ilu_factor_color<2, TM>), dim3(grid_for(nr * B, bs)), dim3(bs), 0, c->stream,
                           (const TM *)im->Ac, (TM *)im->LU, N, im->cnb, im->cnp, im->clow, im->color, im->crow,
                           (TM *)im->Dr, (TM *)im->Ur, r0, nr, lgB);
      else
        hipLaunchKernelGGL((k_ilu_factor_color<3, TM>), dim3(grid_for(nr * B, bs)), dim3(bs), 0, c->stream,
                           (const TM *)im->Ac, (TM *)im->LU, N, im->cnb, im->cnp, im->clow, im->color, im->crow,
                           (TM *)im->Dr, (TM *)im->Ur, r0, nr, lgB);
    }
  }
  BES_HIP(hipGetLastError());
  return BES_OK;
}

// z = M^-1 r (B conditions at once for the block Jacobi and ILU(0) forms)
template <typename TM, typename TV>
static int precond_apply(bes_ctx *c, Impl *im, const TV *r, TV *z, TV *tmp) {
  if (im->precond == 4) return amg_apply<TV>(c, im, r, z, tmp);
  const int bs = tuned_bs(c, kProfPrecond), N = c->N, B = c->B, lgB = c->lgB;
  const size_t n = (size_t)N * c->nv * B;
  const double vb = (double)n * sizeof(TV), blocks = (double)(N + 2 * im->ninternal) * c->nv * c->nv * B;
  // ILU: every stored block once over both sweeps, r, z, y written and read for
  // the colours between the first and the last, neighbour ids and colours;
  // Jacobi: diagonal inverse, r, z
  const int nc = im->ncolor;
  const double ymid = nc > 1 ? (double)(im->color_start[nc - 1] - im->color_start[1]) / std::max(N, 1) : 0.0;
  ProfScope ps(c, kProfPrecond,
               im->precond == 2 ? blocks * sizeof(TM) + (2.0 + 2.0 * ymid) * vb + 2.0 * N * (4.0 * c->k + 4.0)
               : im->precond == 3 ? blocks * (im->prec == 32 ? 2.0 : (double)sizeof(TM)) + 2.0 * vb + 4.0 * N * c->k
               : (im->precond == 1 ? (double)N * c->nv * c->nv * B * sizeof(TM) + 2.0 * vb : 2.0 * vb));
  if (im->precond == 0) {
    hipLaunchKernelGGL((k_scale<TV, TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, r, 1.0, z, n);
  } else if (im->precond == 3 && im->prec == 32) {   // learned operator, FP16 copy: one block SpMV
    if (c->nd == 2) hipLaunchKernelGGL((k_spmv<2, __half, TV>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const __half *)im->LMh, c->dm, r, z, 0);
    else hipLaunchKernelGGL((k_spmv<3, __half, TV>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const __half *)im->LMh, c->dm, r, z, 0);
  } else if (im->precond == 3) {   // learned operator: one block SpMV
    if (c->nd == 2) hipLaunchKernelGGL((k_spmv<2, TM, TV>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const TM *)im->LM, c->dm, r, z, 0);
    else hipLaunchKernelGGL((k_spmv<3, TM, TV>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const TM *)im->LM, c->dm, r, z, 0);
  } else if (im->precond == 1) {
    if (c->nd == 2) hipLaunchKernelGGL((k_bj_apply<2, TM, TV>), dim3(grid_for(N * B, bs)), dim3(bs), 0, c->stream, (const TM *)im->LU, N, r, z, lgB);
    else hipLaunchKernelGGL((k_bj_apply<3, TM, TV>), dim3(grid_for(N * B, bs)), dim3(bs), 0, c->stream, (const TM *)im->LU, N, r, z, lgB);
  } else {
    // colour 0 needs no forward sweep (y = r); the last colour's backward step
    // is fused into its forward sweep
    const int y0end = nc > 0 ? im->color_start[1] : 0;
    for (int col = 1; col < nc; col++) {
      const int r0 = im->color_start[col], nr = im->color_start[col + 1] - r0;
      if (nr == 0) continue;
      const dim3 g(grid_for(nr * B, bs));
      if (col == nc - 1) {
        if (c->nd == 2)
          hipLaunchKernelGGL((k_ilu_lower<2, TM, TV, true>), g, dim3(bs), 0, c->stream, (const TM *)im->LU, N, im->cnp,
                             im->clow, r0, nr, r, tmp, y0end, z, lgB);
        else
          hipLaunchKernelGGL((k_ilu_lower<3, TM, TV, true>), g, dim3(bs), 0, c->stream, (const TM *)im->LU, N, im->cnp,
                             im->clow, r0, nr, r, tmp, y0end, z, lgB);
      } else if (c->nd == 2) {
        hipLaunchKernelGGL((k_ilu_lower<2, TM, TV, false>), g, dim3(bs), 0, c->stream, (const TM *)im->LU, N, im->cnp,
                           im->clow, r0, nr, r, tmp, y0end, (TV *)nullptr, lgB);
      } else {
        hipLaunchKernelGGL((k_ilu_lower<3, TM, TV, false>), g, dim3(bs), 0, c->stream, (const TM *)im->LU, N, im->cnp,
                           im->clow, r0, nr, r, tmp, y0end, (TV *)nullptr, lgB);
      }
    }
    for (int col = nc > 1 ? nc - 2 : nc - 1; col >= 0; col--) {
      const int r0 = im->color_start[col], nr = im->color_start[col + 1] - r0;
      if (nr == 0) continue;
      if (c->nd == 2)
        hipLaunchKernelGGL((k_ilu_upper<2, TM, TV>), dim3(grid_for(nr * B, bs)), dim3(bs), 0, c->stream,
                           (const TM *)im->LU, N, im->cnp, im->clow, r0, nr, tmp, z, r, y0end, lgB);
      else
        hipLaunchKernelGGL((k_ilu_upper<3, TM, TV>), dim3(grid_for(nr * B, bs)), dim3(bs), 0, c->stream,
                           (const TM *)im->LU, N, im->cnp, im->clow, r0, nr, tmp, z, r, y0end, lgB);
    }
  }
  BES_HIP(hipGetLastError());
  return BES_OK;
}

template <typename TM, typename TV>
static int spmv(bes_ctx *c, Impl *im, const TV *x, TV *y) {
  const int bs = tuned_bs(c, kProfSpmv), B = c->B, lgB = c->lgB;
  // stored blocks, x and y once, neighbour ids
  ProfScope ps(c, kProfSpmv, (double)(c->N + 2 * im->ninternal) * c->nv * c->nv * sizeof(TM) * B +
                                 2.0 * c->N * c->nv * sizeof(TV) * B + 4.0 * c->N * c->k);
  if (im->precond == 2) {   // ILU(0) solves run in colour order
    if (c->nd == 2)
      hipLaunchKernelGGL((k_spmv_pos<2, TM, TV>), dim3(grid_for(c->N * B, bs)), dim3(bs), 0, c->stream,
                         (const TM *)im->Ac, im->cnp, c->N, x, y, lgB);
    else
      hipLaunchKernelGGL((k_spmv_pos<3, TM, TV>), dim3(grid_for(c->N * B, bs)), dim3(bs), 0, c->stream,
                         (const TM *)im->Ac, im->cnp, c->N, x, y, lgB);
  } else if (c->nd == 2) {
    hipLaunchKernelGGL((k_spmv<2, TM, TV>), dim3(grid_for(c->N * B, bs)), dim3(bs), 0, c->stream, (const TM *)im->A, c->dm, x, y, lgB);
  } else {
    hipLaunchKernelGGL((k_spmv<3, TM, TV>), dim3(grid_for(c->N * B, bs)), dim3(bs), 0, c->stream, (const TM *)im->A, c->dm, x, y, lgB);
  }
  BES_HIP(hipGetLastError());
  return BES_OK;
}

// the lane sums of rows qb = q*B + b < nqb of S (L lanes each) into
// out[q'*out_stride_q + b], q' = 8 for q = qnorm, else q: level 1 over groups
// of 1024 lanes, then level 2 over the groups (skipped for L <= 1024)
static void sum_lanes(bes_ctx *c, Impl *im, const double *S, int nqb, double *out, int out_stride_q, int qnorm) {
  const int L = im->lanes, G = (L + 1023) / 1024;
  if (G == 1) {
    hipLaunchKernelGGL(k_sum_lanes1, dim3(1, nqb), dim3(256), 0, c->stream, S, L, out, out_stride_q, qnorm, c->lgB);
  } else {
    hipLaunchKernelGGL(k_sum_lanes1, dim3(G, nqb), dim3(256), 0, c->stream, S, L, im->S2, out_stride_q, qnorm,
                       c->lgB);
    hipLaunchKernelGGL(k_sum_lanes, dim3(nqb), dim3(256), 0, c->stream, im->S2, G, out, out_stride_q, qnorm, c->lgB);
  }
}

// dots h_q = V_q^T w (q < cnt) and ||w||^2 of every condition: chunk ch of up to
// 8 vectors reduces into dots[(9*ch + q)*B + b], its ||w||^2 into
// dots[(9*ch + 8)*B + b] (the deterministic lane reductions)
template <typename TV>
static int multidot(bes_ctx *c, Impl *im, const TV *V, size_t stride, int cnt, const TV *w) {
  const int B = c->B, lgB = c->lgB, L = im->lanes;
  const size_t n1 = (size_t)c->N * c->nv;
  const int nchunk = std::max(1, (cnt + 7) / 8);
  if (nchunk > 4) { set_error("too many basis vectors"); return BES_ERR_ARG; }
  ProfScope ps(c, kProfKrylov, (double)(cnt + nchunk) * n1 * B * sizeof(TV));
  const int nt = L * B, nb = grid_for(nt, 256);
  for (int q = 0; q < nchunk; q++) {
    const int i0 = q * 8, cc = std::max(0, std::min(8, cnt - i0));
    hipLaunchKernelGGL((k_mdot<TV>), dim3(nb), dim3(256), 0, c->stream, V, stride, i0, cc, w, n1, lgB, L, im->S);
    sum_lanes(c, im, im->S, (cc + 1) * B, im->dots + (size_t)9 * q * B, B, cc);
  }
  BES_HIP(hipGetLastError());
  return BES_OK;
}

// one lane-reduced norm per condition (S slot 0) into dots[at*B + b]
static int sum_norms(bes_ctx *c, Impl *im, int at) {
  sum_lanes(c, im, im->S, c->B, im->dots + (size_t)at * c->B, c->B, -1);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

static int fetch_dots(bes_ctx *c, Impl *im, size_t ndbl) {
  BES_HIP(hipMemcpyAsync(im->hdots, im->dots, ndbl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  return BES_OK;
}

// GMRES(m), right-preconditioned, for the B conditions of the batch at once
// (one launch per kernel for all of them).  Each condition keeps its own
// Arnoldi process, Givens rotations, iteration count and stopping point: a
// condition whose estimate meets its target, whose basis breaks down or that
// reaches maxiter stops at that step of the cycle (its x uses exactly those
// steps), the cycle goes on while any condition is active, and every cycle
// restarts from the true residual of each condition still unconverged.  All
// reductions are the lane reductions above, so a condition's iterates are the
// same bits in a batch as alone.  Classical Gram-Schmidt; with eta > 0 a DGKS
// second pass for the conditions whose w kept less than eta of its norm.
template <typename TM, typename TV>
static int gmres(bes_ctx *c, Impl *im, int *iters_out, double *relres_out) {
  const int m = im->restart, B = c->B, lgB = c->lgB;
  const size_t n1 = (size_t)c->N * c->nv, n = n1 * B, vs = im->vstride;
  const int L = im->lanes, nbl = grid_for(L * B, 256);
  TV *V = (TV *)im->V;
  TV *Z = (TV *)im->Z, *Z1 = Z + vs, *Z2 = Z + 2 * vs;
  // ILU(0) solves run in colour order: b in, x out are permuted once per solve
  const bool colour = im->precond == 2;
  const double *bb = im->b;
  double *xx = colour ? im->xc : im->x;
  if (colour) {
    hipLaunchKernelGGL(k_perm_rows<double>, dim3(1024), dim3(256), 0, c->stream, bb, im->crow, c->N, c->nv * B, 1,
                       im->bc);
    bb = im->bc;
  }
  hipMemsetAsync(xx, 0, n * sizeof(double), c->stream);
  const double *hd = im->hdots;
  auto hdot = [&](int q, int b) { return hd[(size_t)(9 * (q / 8) + (q % 8)) * B + b]; };
  // ||b|| per condition
  hipLaunchKernelGGL((k_scale<double, TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, bb, 1.0, Z, n);
  int s;
  if ((s = multidot<TV>(c, im, Z, vs, 0, Z)) || (s = fetch_dots(c, im, (size_t)9 * B))) return s;
  std::vector<double> bnorm(B), target(B), rnorm(B);
  std::vector<char> finished(B, 0), active(B, 0), est_done(B, 0);
  std::vector<int> total(B, 0), jb(B, 0);
  for (int b = 0; b < B; b++) {
    bnorm[b] = std::sqrt(hd[(size_t)8 * B + b]);
    target[b] = im->rtol * bnorm[b];
    rnorm[b] = bnorm[b];
    finished[b] = bnorm[b] == 0 || (!im->skip_b.empty() && im->skip_b[b]);
  }
  std::vector<double> H((size_t)B * (m + 1) * m), cs((size_t)B * m), sn((size_t)B * m), g((size_t)B * (m + 1)),
      yv((size_t)B * m), hcol(m + 1), coef((size_t)m * B);
  CoefB scl;
  bool first = true;
  auto any_open = [&]() {
    for (int b = 0; b < B; b++) if (!finished[b] && total[b] < im->maxiter) return true;
    return false;
  };
  while (any_open()) {
    // r = b - A x (x = 0 on the first cycle, as the reference starts every
    // solve).  Every later cycle starts from the true residual of the current
    // x, and a condition is done only when that is at its target: the Givens
    // estimate of the cycle before can fall below the true residual (rounding
    // in the basis, FP32 Krylov vectors), as AMGCL's FGMRES also rechecks
    // (fgmres.hpp:158-163)
    if (first) {
      for (int b = 0; b < B; b++) scl.c[b] = finished[b] ? 0.0 : 1.0 / bnorm[b];
      hipLaunchKernelGGL((k_scale_b<double, TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, bb, scl, V, n, lgB);
    } else {
      hipLaunchKernelGGL((k_scale<double, TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, xx, 1.0, Z2, n);
      if ((s = spmv<TM, TV>(c, im, Z2, Z1))) return s;
      hipLaunchKernelGGL((k_resid<TV>), dim3(nbl), dim3(256), 0, c->stream, bb, Z1, Z, n1, lgB, L, im->S);
      if ((s = sum_norms(c, im, 0)) || (s = fetch_dots(c, im, (size_t)B))) return s;
      for (int b = 0; b < B; b++) {
        if (finished[b]) continue;
        rnorm[b] = std::sqrt(hd[b]);
        if (rnorm[b] <= target[b]) finished[b] = 1;
        else if (est_done[b]) im->extra_restarts++;   // the estimate said done, the true residual did not
      }
      if (!any_open()) break;
      for (int b = 0; b < B; b++) scl.c[b] = (!finished[b] && total[b] < im->maxiter) ? 1.0 / rnorm[b] : 0.0;
      hipLaunchKernelGGL((k_scale_b<TV, TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, Z, scl, V, n, lgB);
    }
    first = false;
    int nact = 0;
    for (int b = 0; b < B; b++) {
      active[b] = !finished[b] && total[b] < im->maxiter;
      est_done[b] = 0;
      jb[b] = 0;
      nact += active[b];
      std::fill(g.begin() + (size_t)b * (m + 1), g.begin() + (size_t)(b + 1) * (m + 1), 0.0);
      g[(size_t)b * (m + 1)] = rnorm[b];
    }
    bool queued = false;
    for (int j = 0; j < m && nact > 0; j++) {
      TV *vj = V + (size_t)j * vs, *w = V + (size_t)(j + 1) * vs;
      if (!queued) {
        if ((s = precond_apply<TM, TV>(c, im, vj, Z, Z1))) return s;
        if ((s = spmv<TM, TV>(c, im, Z, w))) return s;
      }
      queued = false;
      // classical Gram-Schmidt: the dots, the update and (without
      // re-orthogonalisation) the normalisation are queued back to back and
      // the host waits once, for h and ||w||
      const bool dev_scale = !(im->reorth_eta > 0);
      if ((s = multidot<TV>(c, im, V, vs, j + 1, w))) return s;
      {
        ProfScope ps(c, kProfKrylov, (double)(j + 3) * n * sizeof(TV));
        hipLaunchKernelGGL((k_orth_update<TV>), dim3(nbl), dim3(256), 0, c->stream, V, vs, j + 1, im->dots, w, n1, lgB,
                           L, im->S);
        if ((s = sum_norms(c, im, 40))) return s;
      }
      if (dev_scale)
        hipLaunchKernelGGL((k_scale_norm<TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, w, im->dots + (size_t)40 * B,
                           n, lgB);
      BES_HIP(hipMemcpyAsync(im->hdots, im->dots, (size_t)41 * B * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      BES_HIP(hipEventRecord(im->ev_dots, c->stream));
      // the next step's preconditioner and product need only the normalised w:
      // queue them before waiting, so the device works while the host reads h
      // and ||w|| (their results are discarded for conditions this step ends)
      bool more = false;
      for (int b = 0; b < B; b++) if (active[b] && total[b] + 1 < im->maxiter) more = true;
      if (dev_scale && j + 1 < m && more) {
        if ((s = precond_apply<TM, TV>(c, im, w, Z, Z1))) return s;
        if ((s = spmv<TM, TV>(c, im, Z, V + (size_t)(j + 2) * vs))) return s;
        queued = true;
      }
      BES_HIP(hipEventSynchronize(im->ev_dots));
      std::vector<double> wn2(B), w0n2(B);
      for (int b = 0; b < B; b++) { w0n2[b] = hd[(size_t)8 * B + b]; wn2[b] = hd[(size_t)40 * B + b]; }
      std::vector<double> hsave;
      if (!dev_scale) {   // DGKS second pass for the conditions that need it (h2 = 0 for the rest: w unchanged)
        bool any = false;
        std::vector<char> re(B, 0);
        for (int b = 0; b < B; b++)
          if (active[b] && std::sqrt(wn2[b]) < im->reorth_eta * std::sqrt(w0n2[b])) { re[b] = 1; any = true; }
        hsave.assign(hd, hd + (size_t)36 * B);
        if (any) {
          if ((s = multidot<TV>(c, im, V, vs, j + 1, w)) || (s = fetch_dots(c, im, (size_t)36 * B))) return s;
          std::vector<double> h2(hd, hd + (size_t)36 * B);
          for (int b = 0; b < B; b++)
            for (int q = 0; q <= j; q++) {
              const size_t at = (size_t)(9 * (q / 8) + (q % 8)) * B + b;
              if (!re[b]) h2[at] = 0.0;
            }
          BES_HIP(hipMemcpyAsync(im->dots, h2.data(), h2.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
          for (int b = 0; b < B; b++)
            if (re[b]) {
              im->reorths++;
              for (int q = 0; q <= j; q++) {
                const size_t at = (size_t)(9 * (q / 8) + (q % 8)) * B + b;
                hsave[at] += h2[at];
              }
            }
          {
            ProfScope ps(c, kProfKrylov, (double)(j + 3) * n * sizeof(TV));
            hipLaunchKernelGGL((k_orth_update<TV>), dim3(nbl), dim3(256), 0, c->stream, V, vs, j + 1, im->dots, w, n1,
                               lgB, L, im->S);
            if ((s = sum_norms(c, im, 40))) return s;
          }
          BES_HIP(hipMemcpyAsync(im->hdots + (size_t)40 * B, im->dots + (size_t)40 * B, B * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
          BES_HIP(hipStreamSynchronize(c->stream));
          for (int b = 0; b < B; b++) wn2[b] = hd[(size_t)40 * B + b];
        }
        for (int b = 0; b < B; b++) {
          const double hn = std::sqrt(wn2[b]);
          scl.c[b] = (active[b] && hn > 0) ? 1.0 / hn : 1.0;
        }
        hipLaunchKernelGGL((k_scale_b<TV, TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, w, scl, w, n, lgB);
      }
      for (int b = 0; b < B; b++) {
        if (!active[b]) continue;
        for (int q = 0; q <= j; q++)
          hcol[q] = dev_scale ? hdot(q, b) : hsave[(size_t)(9 * (q / 8) + (q % 8)) * B + b];
        const double hn = std::sqrt(wn2[b]);
        hcol[j + 1] = hn;
        double *csb = &cs[(size_t)b * m], *snb = &sn[(size_t)b * m], *gb = &g[(size_t)b * (m + 1)];
        for (int q = 0; q < j; q++) {   // Givens rotations
          const double t = csb[q] * hcol[q] + snb[q] * hcol[q + 1];
          hcol[q + 1] = -snb[q] * hcol[q] + csb[q] * hcol[q + 1];
          hcol[q] = t;
        }
        const double den = std::hypot(hcol[j], hcol[j + 1]);
        csb[j] = den > 0 ? hcol[j] / den : 1.0;
        snb[j] = den > 0 ? hcol[j + 1] / den : 0.0;
        hcol[j] = den;
        hcol[j + 1] = 0;
        gb[j + 1] = -snb[j] * gb[j];
        gb[j] = csb[j] * gb[j];
        for (int q = 0; q <= j; q++) H[((size_t)b * (m + 1) + q) * m + j] = hcol[q];
        total[b]++;
        rnorm[b] = std::fabs(gb[j + 1]);
        jb[b] = j + 1;
        if (rnorm[b] <= target[b] || hn == 0) { est_done[b] = 1; active[b] = 0; nact--; }
        else if (total[b] >= im->maxiter) { active[b] = 0; nact--; }
      }
    }
    // solve H y = g (upper triangular j_b x j_b) per condition; x += M (V y)
    int jmax = 0;
    for (int b = 0; b < B; b++) jmax = std::max(jmax, jb[b]);
    if (jmax == 0) continue;
    std::fill(coef.begin(), coef.end(), 0.0);
    for (int b = 0; b < B; b++) {
      const int j = jb[b];
      const double *gb = &g[(size_t)b * (m + 1)];
      double *yb = &yv[(size_t)b * m];
      for (int q = j - 1; q >= 0; q--) {
        double t = gb[q];
        for (int p = q + 1; p < j; p++) t -= H[((size_t)b * (m + 1) + q) * m + p] * yb[p];
        yb[q] = t / H[((size_t)b * (m + 1) + q) * m + q];
      }
      for (int q = 0; q < j; q++) coef[(size_t)q * B + b] = yb[q];
    }
    BES_HIP(hipMemcpyAsync(im->dots, coef.data(), sizeof(double) * (size_t)jmax * B, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL((k_combine<TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, V, vs, jmax, im->dots, Z2, n, lgB);
    if ((s = precond_apply<TM, TV>(c, im, Z2, Z, Z1))) return s;
    hipLaunchKernelGGL((k_axpy_x<TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, xx, Z, n);
    BES_HIP(hipGetLastError());
  }
  // true residual of the returned solution, per condition
  hipLaunchKernelGGL((k_scale<double, TV>), dim3(kRedBlocks), dim3(256), 0, c->stream, xx, 1.0, Z2, n);
  if ((s = spmv<TM, TV>(c, im, Z2, Z1))) return s;
  hipLaunchKernelGGL((k_resid<TV>), dim3(nbl), dim3(256), 0, c->stream, bb, Z1, Z, n1, lgB, L, im->S);
  if ((s = sum_norms(c, im, 0)) || (s = fetch_dots(c, im, (size_t)B))) return s;
  if (colour)
    hipLaunchKernelGGL(k_perm_rows<double>, dim3(1024), dim3(256), 0, c->stream, im->xc, im->crow, c->N, c->nv * B, 0,
                       im->x);
  im->iters_b.assign(total.begin(), total.end());
  im->relres_b.assign(B, 0.0);
  int itmax = 0;
  double rrmax = 0;
  for (int b = 0; b < B; b++) {
    im->relres_b[b] = bnorm[b] > 0 ? std::sqrt(hd[b]) / bnorm[b] : 0.0;
    if (!im->skip_b.empty() && im->skip_b[b]) im->relres_b[b] = 0.0;
    itmax = std::max(itmax, total[b]);
    rrmax = std::max(rrmax, im->relres_b[b]);
  }
  *iters_out = itmax;
  *relres_out = rrmax;
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
  hipLaunchKernelGGL((k_resid<TV>), dim3(grid_for(im->lanes, 256)), B, 0, c->stream, bb, Z1, Z, n, 0, im->lanes,
                     im->S);
  if ((st = sum_norms(c, im, 0)) || (st = fetch_dots(c, im, 1))) return st;
  d[0] = im->hdots[0];
  if (colour)
    hipLaunchKernelGGL(k_perm_rows<double>, dim3(1024), B, 0, c->stream, im->xc, im->crow, c->N, c->nv, 0, im->x);
  BES_HIP(hipGetLastError());
  *iters_out = iter;
  *relres_out = std::sqrt(d[0]) / bnorm;
  return BES_OK;
}

// average device time of fn() over reps launches on the context's stream (one untimed first)
template <typename Fn>
static int time_us(bes_ctx *c, int reps, Fn fn, double *us) {
  hipEvent_t a, b;
  BES_HIP(hipEventCreate(&a));
  BES_HIP(hipEventCreate(&b));
  int s = fn();
  if (!s) {
    hipEventRecord(a, c->stream);
    for (int r = 0; r < reps && !s; r++) s = fn();
    hipEventRecord(b, c->stream);
    hipEventSynchronize(b);
    float ms = 0;
    hipEventElapsedTime(&ms, a, b);
    *us = 1e3 * ms / reps;
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  return s;
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
  if (c->B != 1 && precond > 2) {
    set_error("a batch of conditions runs with precond 0-2 (none, block Jacobi, ILU(0)); the learned operator and "
              "the AMG take one condition");
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

int bes_implicit_set_coloring(bes_ctx *c, int mode) {
  if (!c || !c->impl || mode < 0 || mode > 2) { set_error("call bes_implicit_setup first / coloring 0-2"); return BES_ERR_ARG; }
  Impl *im = impl_of(c);
  if (im->coloring == mode) return BES_OK;
  BES_HIP(hipStreamSynchronize(c->stream));
  return build_colouring(c, im, mode);
}

int bes_implicit_ncolors(bes_ctx *c) {
  if (!c || !c->impl) { set_error("call bes_implicit_setup first"); return -BES_ERR_ARG; }
  return impl_of(c)->ncolor;
}

int bes_implicit_set_solver(bes_ctx *c, int solver) {
  if (!c || !c->impl) { set_error("call bes_implicit_setup first"); return BES_ERR_ARG; }
  if (solver != BES_SOLVER_GMRES && solver != BES_SOLVER_BICGSTAB) {
    set_error("LINEAR_SOLVER: only GMRES (FGMRES) and BICGSTAB are available; LGMRES is not");
    return BES_ERR_ARG;
  }
  if (solver == BES_SOLVER_BICGSTAB && c->B != 1) {
    set_error("LINEAR_SOLVER: BICGSTAB takes one condition; a batch runs GMRES");
    return BES_ERR_ARG;
  }
  impl_of(c)->solver = solver;
  return BES_OK;
}

int bes_implicit_set_precond(bes_ctx *c, int precond) {
  if (!c || !c->impl || precond < 0 || precond > 4) { set_error("call bes_implicit_setup first / precond 0-4"); return BES_ERR_ARG; }
  if (c->B != 1 && precond > 2) { set_error("a batch of conditions runs with precond 0-2"); return BES_ERR_ARG; }
  Impl *im = impl_of(c);
  im->precond = precond;
  im->factored = false;
  return BES_OK;
}

// time step, gradient, limiter, Jacobians and system of every condition; cfl_b
// (per condition) may be null; linf outputs [B][nv]
static int assemble(bes_ctx *c, Impl *im, double cfl, const double *cfl_b, int local, int repeat_nv, double *linf_b,
                    double *linf_r) {
  im->repeat_nv = repeat_nv ? c->nv : 1;
  const int B = c->B;
  int s;
  hipLaunchKernelGGL(k_fill_u64_i, dim3(1), dim3(64), 0, c->stream, c->err, 1, ~0ull);
  if ((s = launch_timestep(c, c->U0, cfl, local, cfl_b))) return s;
  c->dt_local = local != 0;
  if ((s = launch_gradlim(c, c->U0, nullptr, nullptr))) return s;
  if (im->prec == 64) s = launch_assemble<double>(c, im);
  else s = launch_assemble<float>(c, im);
  if (s) return s;
  hipLaunchKernelGGL(k_linf_reduce_i, dim3(grid_for(16 * B, 64)), dim3(64), 0, c->stream, im->bits, 16 * B);
  std::vector<unsigned long long> bb((size_t)16 * B);
  BES_HIP(hipMemcpyAsync(bb.data(), im->bits, bb.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                         c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  for (int b = 0; b < B; b++)
    for (int v = 0; v < c->nv; v++) {   // bits [v][B] of b, then [v][B] of sum F*A
      if (linf_b) std::memcpy(&linf_b[(size_t)b * c->nv + v], &bb[(size_t)v * B + b], 8);
      if (linf_r) std::memcpy(&linf_r[(size_t)b * c->nv + v], &bb[(size_t)(8 + v) * B + b], 8);
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

int bes_implicit_assemble(bes_ctx *c, double cfl, int local, int repeat_nv, double *linf_b, double *linf_r) {
  if (!c || !c->impl) { set_error("call bes_implicit_setup first"); return BES_ERR_ARG; }
  return assemble(c, impl_of(c), cfl, nullptr, local, repeat_nv, linf_b, linf_r);
}

// reference CoefficientMatrix layout: scalar CSR rows (cell, var) in reference
// cell order, each row = own block row then neighbour block rows in local face
// order (CoefficientMatrix.cpp:127-160).  rowptr [ncell*nv+1], col/val [nnz].
int bes_implicit_get_system(bes_ctx *c, long *rowptr, long *col, double *val, double *rhs) {
  if (!c || !c->impl || !impl_of(c)->assembled) { set_error("system not assembled"); return BES_ERR_ARG; }
  if (c->B != 1) { set_error("the CSR system is available for one condition (batch size 1)"); return BES_ERR_ARG; }
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
  if (c->B != 1) { set_error("a CSR system is loaded for one condition (batch size 1)"); return BES_ERR_ARG; }
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
  if (im->precond == 3 && im->lp_online_steps > 0) {   // online training on the system just assembled
    std::vector<double> lh(im->lp_online_steps);
    if ((s = bes_lp_train(c, im->lp_online_steps, im->lp_online_lr, im->lp_online_nprobe, im->lp_online_nrich,
                          im->lp_online_seed, lh.data())))
      return s;
    im->lp_online_seed += (unsigned long long)im->lp_online_steps;
    im->lp_online_loss = lh.back();
  }
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
  const int N = c->N, nv = c->nv, B = c->B;
  std::vector<double> h((size_t)N * nv * B);
  BES_HIP(hipMemcpy(h.data(), im->x, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int b = 0; b < B; b++)
    for (int64_t e = 0; e < N; e++)
      for (int v = 0; v < nv; v++)
        x[((size_t)b * N + e) * nv + v] = h[((size_t)c->iperm[e] * nv + v) * B + b];
  return BES_OK;
}

int bes_amg_apply(bes_ctx *c, const double *r, double *z) {
  if (!c || !c->impl || !r || !z) { set_error("call bes_implicit_setup first"); return BES_ERR_ARG; }
  if (c->B != 1) { set_error("the AMG takes one condition"); return BES_ERR_ARG; }
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

// U += x for the conditions whose active flag is set (all when active is null)
static int update(bes_ctx *c, Impl *im, const std::vector<char> *active) {
  const int bs = tuned_bs(c, kProfUpdate), B = c->B;
  CoefB act;
  for (int b = 0; b < B; b++) act.c[b] = (!active || (*active)[b]) ? 1.0 : 0.0;
  {
    ProfScope ps(c, kProfUpdate, 24.0 * c->N * c->nv * B);
    if (c->nd == 2)
      hipLaunchKernelGGL(k_update<2>, dim3(grid_for(c->N * B, bs)), dim3(bs), 0, c->stream, c->U0, im->x, c->N, c->lgB,
                         act);
    else
      hipLaunchKernelGGL(k_update<3>, dim3(grid_for(c->N * B, bs)), dim3(bs), 0, c->stream, c->U0, im->x, c->N, c->lgB,
                         act);
  }
  BES_HIP(hipGetLastError());
  BES_HIP(hipStreamSynchronize(c->stream));
  return prof_collect(c);
}

int bes_implicit_update(bes_ctx *c) {
  if (!c || !c->impl) { set_error("call bes_implicit_setup first"); return BES_ERR_ARG; }
  return update(c, impl_of(c), nullptr);
}

// SolverImplicit::solve loop (SolverImplicit.cpp:552-600) with the adaptive
// Courant number of SolverImplicit.cpp:530-550, for every condition of the
// batch at once: each keeps its own Courant number, residual history and stop;
// a condition that has reached its drop is frozen (no solve, no update) while
// the others go on.  hist [nsteps][B][2*nv+3]: Linf(b) per var, Linf(sum F*A)
// per var, CFL used, Krylov iterations, relres (NaN rows for frozen conditions).
static int run_steps(bes_ctx *c, Impl *im, int nsteps, double *hist, int *steps_done) {
  const int nv = c->nv, B = c->B, hw = 2 * nv + 3;
  Impl::Run &R = im->run;
  std::vector<double> lb((size_t)B * nv), lr((size_t)B * nv);
  std::vector<char> go(B);
  int done = 0;
  for (int it = 0; it < nsteps && !R.reached; it++) {
    for (int b = 0; b < B; b++) go[b] = !R.done[b];
    im->skip_b.assign(R.done.begin(), R.done.end());
    int s = assemble(c, im, 0.0, R.cfl.data(), R.local, R.repeat_nv, lb.data(), lr.data());
    if (s) { im->skip_b.clear(); if (steps_done) *steps_done = done; return s; }
    int iters = 0;
    double rr = 0;
    s = bes_linear_solve(c, &iters, &rr);
    im->skip_b.clear();
    if (s && s != BES_ERR_SOLVE) { if (steps_done) *steps_done = done; return s; }
    if ((s = update(c, im, &go))) { if (steps_done) *steps_done = done; return s; }
    done = it + 1;
    bool all = true;
    for (int b = 0; b < B; b++) {
      const double *lbb = &lb[(size_t)b * nv], *lrb = &lr[(size_t)b * nv];
      if (hist) {
        double *h = hist + ((size_t)it * B + b) * hw;
        if (!go[b]) {
          for (int q = 0; q < hw; q++) h[q] = std::nan("");
        } else {
          for (int v = 0; v < nv; v++) { h[v] = lbb[v]; h[nv + v] = lrb[v]; }
          h[2 * nv] = R.cfl[b];
          h[2 * nv + 1] = B == 1 ? iters : im->iters_b[b];
          h[2 * nv + 2] = B == 1 ? rr : im->relres_b[b];
        }
      }
      if (go[b]) {
        R.steps[b]++;
        if (R.r0[b] < 0) R.r0[b] = lrb[0];
        // limiter freeze: from the step whose density residual first reaches
        // the freeze drop on, the condition keeps the limiter of that step's
        // gradient pass (stored by it in phi_frz)
        if (c->lim_freeze_drop > 0 && b < (int)c->lim_frozen.size() && !c->lim_frozen[b] && R.r0[b] > 0 &&
            lrb[0] <= c->lim_freeze_drop * R.r0[b]) {
          c->lim_frozen[b] = 1;
          c->lim_epoch++;
        }
        // adaptive Courant number (density monitor of the implicit residual)
        if (R.cfl_max > R.cfl_min) {
          double &lo = R.lb_old[(size_t)b * nv];
          const int ei = (lo > lbb[0]) ? 1 : 0;
          double cf = R.cfl[b] * std::pow(lo / lbb[0], R.ex[ei]);
          cf = std::max(R.cfl_min, cf);
          R.cfl[b] = std::min(R.cfl_max, cf);
        }
        std::copy(lbb, lbb + nv, R.lb_old.begin() + (size_t)b * nv);
        if (R.stop_drop > 0 && R.r0[b] > 0 && lrb[0] <= R.stop_drop * R.r0[b]) R.done[b] = 1;
      }
      all = all && R.done[b];
    }
    R.reached = all;
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
  R.cfl_min = cfl_min;
  R.cfl.assign(c->B, cfl_min);
  R.r0.assign(c->B, -1.0);
  R.done.assign(c->B, 0);
  R.steps.assign(c->B, 0);
  R.cfl_max = cfl_max;
  R.ex[0] = exponents[0];
  R.ex[1] = exponents[1];
  R.local = local;
  R.repeat_nv = repeat_nv;
  R.stop_drop = stop_drop;
  R.lb_old.assign((size_t)c->B * c->nv, 0.0);
  return run_steps(c, im, nsteps, hist, steps_done);
}

int bes_autotune(bes_ctx *c, int flags, int *chosen, double *best_us, int *coloring) {
  if (!c || flags < 0 || flags > 7) { set_error("invalid autotune request (flags 0-7)"); return BES_ERR_ARG; }
  Impl *im = (Impl *)c->impl;
  if ((flags & 6) && !(im && im->assembled)) { set_error("implicit autotuning needs an assembled system"); return BES_ERR_ARG; }
  if ((flags & 6) && (im->precond == 3 || im->precond == 4)) {
    set_error("implicit autotuning covers precond 0-2");
    return BES_ERR_ARG;
  }
  BES_HIP(hipStreamSynchronize(c->stream));
  const bool prof = c->prof;
  c->prof = false;
  if (c->graph) { hipGraphExecDestroy(c->graph); c->graph = nullptr; }   // launches captured with old sizes
  const int cand[3] = {64, 128, 256};
  double best[bes_ctx::kProfKinds];
  for (int k = 0; k < bes_ctx::kProfKinds; k++) best[k] = 0;
  int s = BES_OK;
  auto tune = [&](int kind, auto fn) -> int {
    double bt = 1e300;
    int bb = 256;
    for (int bs : cand) {
      c->tune_bs[kind] = bs;
      double us = 0;
      int st = time_us(c, 5, fn, &us);
      if (st) return st;
      if (us < bt) { bt = us; bb = bs; }
    }
    c->tune_bs[kind] = bb;
    best[kind] = bt;
    return BES_OK;
  };
  if (!s && (flags & 2)) {   // implicit kernels on the assembled system (re-assembled, re-factored: same values)
    const bool p64 = im->prec == 64;
    auto asmb = [&]() { return p64 ? launch_assemble<double>(c, im) : launch_assemble<float>(c, im); };
    auto fact = [&]() { return p64 ? factor<double>(c, im) : factor<float>(c, im); };
    s = tune(kProfJacobian, asmb);
    if (!s) s = tune(kProfAssemble, asmb);
    if (!s) s = tune(kProfFactor, fact);
    if (!s) {
      const size_t vs = im->vstride;
      if (p64) {
        double *V = (double *)im->V, *Z = (double *)im->Z;
        s = tune(kProfSpmv, [&]() { return spmv<double, double>(c, im, V, V + vs); });
        if (!s) s = tune(kProfPrecond, [&]() { return precond_apply<double, double>(c, im, V, Z, Z + vs); });
      } else {
        float *V = (float *)im->V, *Z = (float *)im->Z;
        s = tune(kProfSpmv, [&]() { return spmv<float, float>(c, im, V, V + vs); });
        if (!s) s = tune(kProfPrecond, [&]() { return precond_apply<float, float>(c, im, V, Z, Z + vs); });
      }
    }
    im->factored = false;
  }
  if (!s && (flags & 4) && im->precond == 2) {   // ILU(0) colouring: the fastest solve of this system
    double bt = 1e300;
    int bm = im->coloring;
    for (int mode = 0; mode <= 2 && !s; mode++) {
      if ((s = build_colouring(c, im, mode))) break;
      int it = 0;
      double rr = 0;
      auto t0 = std::chrono::steady_clock::now();
      s = solve_once(c, im, &it, &rr);
      BES_HIP(hipStreamSynchronize(c->stream));
      const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
      if (us < bt) { bt = us; bm = mode; }
    }
    if (!s) s = build_colouring(c, im, bm);
    if (coloring) *coloring = bm;
  }
  if (!s && (flags & 1)) {   // explicit kernels on the current state, last: the time step it leaves is at CFL 1
    // (U0 is read, never written; the next assembly or explicit step recomputes dt)
    s = tune(kProfGradlim, [&]() { return launch_gradlim(c, c->U0, nullptr, nullptr); });
    if (!s) s = tune(kProfTimestep, [&]() { return launch_timestep(c, c->U0, 1.0, 1); });
    if (!s) s = tune(kProfFaceFlux, [&]() { return launch_face_flux(c); });
    if (!s) s = tune(kProfResidual, [&]() { return launch_residual(c, nullptr, nullptr, 0.0, 1, nullptr, nullptr); });
  }
  if (!s) BES_HIP(hipStreamSynchronize(c->stream));
  c->prof = prof;
  if (chosen)
    for (int k = 0; k < bes_ctx::kProfKinds; k++) chosen[k] = tuned_bs(c, k);
  if (best_us)
    for (int k = 0; k < bes_ctx::kProfKinds; k++) best_us[k] = best[k];
  return s;
}

int bes_implicit_batch_state(bes_ctx *c, int *steps, int *reached, double *cfl, int *iters, double *relres) {
  if (!c || !c->impl) { set_error("call bes_implicit_setup first"); return BES_ERR_ARG; }
  Impl *im = impl_of(c);
  const Impl::Run &R = im->run;
  for (int b = 0; b < c->B; b++) {
    if (steps) steps[b] = R.active ? R.steps[b] : 0;
    if (reached) reached[b] = R.active ? R.done[b] : 0;
    if (cfl) cfl[b] = R.active ? R.cfl[b] : 0.0;
    if (iters) iters[b] = b < (int)im->iters_b.size() ? im->iters_b[b] : 0;
    if (relres) relres[b] = b < (int)im->relres_b.size() ? im->relres_b[b] : 0.0;
  }
  return BES_OK;
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
