// Explicit residual path of the BES solver core on MI355X (gfx950).
//
// Per pseudo-time step (SolverExplicit.cpp:259-393):
//   k_timestep   local dt per cell + global minimum (TimeStepCalculator*.cpp)
//   per stage:
//     k_gradlim  Green-Gauss gradient fused with the Venkatakrishnan limiter,
//                writes the limited gradient phi*grad (GradientGreenGauss.cpp:21-135,
//                LimiterVenkatakrishnan.cpp:21-128)
//     k_residual second-order face fluxes gathered per cell, summed in the
//                reference face order, fused with the stage update
//                U = U0 - alpha*dt/V*R and the Linf monitor
//                (SolverExplicit.cpp:285-378, Solver.cpp:69-121)
// One thread per cell; every cell computes the fluxes of its own faces, so no
// atomics touch the residual and the result is deterministic.  Blocks are
// remapped so that each XCD walks a contiguous range of Morton-ordered cells.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/bes_hip.h"
#include "context.hpp"
#include "flux.hpp"
#include "kernels_common.hpp"

namespace bes {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
int hip_fail(hipError_t e, const char *what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return BES_ERR_HIP;
}

static hipEvent_t take_event(bes_ctx *c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

ProfScope::ProfScope(bes_ctx *c_, int kind_, double bytes_) : c(c_), kind(kind_), bytes(bytes_) {
  if (!c->prof) return;
  a = take_event(c);
  b = take_event(c);
  if (a) hipEventRecord(a, c->stream);
}
ProfScope::~ProfScope() {
  if (!c->prof || !a || !b) return;
  hipEventRecord(b, c->stream);
  c->ev_pending.push_back({kind, bytes, a, b});
}

int prof_collect(bes_ctx *c) {
  if (c->ev_pending.empty()) return BES_OK;
  BES_HIP(hipStreamSynchronize(c->stream));
  for (auto &p : c->ev_pending) {
    float ms = 0;
    BES_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    c->prof_ms[p.kind] += ms;
    c->prof_n[p.kind] += 1;
    c->prof_bytes[p.kind] += p.bytes;
    c->ev_pool.push_back(p.a);
    c->ev_pool.push_back(p.b);
  }
  c->ev_pending.clear();
  return BES_OK;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Batched layout: B freestream conditions share the mesh (B a power of two
// <= 64, lgB = log2 B).  A state value of variable v, cell i, condition b sits
// at ((v*N + i)*B + b): for B = 64 a wave holds one cell's 64 conditions, so
// geometry loads are wave-uniform and state loads are coalesced 512-B rows;
// for B = 1 the layout is the plain SoA U[v][N].

// cfl_b: per-condition Courant numbers when per_cond is set, else cfl for all
// local time step of (cell i, condition b), stored in dt and returned
template <int ND>
__device__ __forceinline__ double timestep_cell(const double *__restrict__ U, const DevMesh &dm, const GasConst &gc,
                                                double cfl, double *__restrict__ dt, int i, int b, int B) {
  constexpr int NV = ND + 2;
  const int N = dm.N;
  double u[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) u[v] = U[((size_t)v * N + i) * B + b];
  Gas<ND, double> g(gc);
  double c = g.sound_speed(u);
  double sr = 0.0;
#pragma unroll
  for (int d = 0; d < ND; d++) sr += (c + ::fabs(g.velocity(u, d))) * dm.sproj[(size_t)d * N + i];
  const double tv = cfl * dm.vol[i] / sr;
  dt[(size_t)i * B + b] = tv;
  return tv;
}

template <int ND>
__global__ __launch_bounds__(256) void k_timestep(const double *__restrict__ U, DevMesh dm, GasConst gc,
                                                  double cfl, CoefB cfl_b, int per_cond, double *__restrict__ dt,
                                                  unsigned long long *__restrict__ gmin, int lgB) {
  const int N = dm.N, B = 1 << lgB;
  const int t = cell_index();
  const int i = t >> lgB, b = t & (B - 1);
  double tv = __builtin_inf();
  if (i < N) tv = timestep_cell<ND>(U, dm, gc, per_cond ? cfl_b.c[b] : cfl, dt, i, b, B);
  if (gmin) {
    double m = block_min_cond(tv, B);
    if ((int)threadIdx.x < B) atomicMin(&gmin[threadIdx.x], (unsigned long long)__double_as_longlong(m));
  }
}

__global__ void k_decode_min(const unsigned long long *bits, double *out, int n) {
  const int q = threadIdx.x;
  if (q < n) out[q] = __longlong_as_double((long long)bits[q]);
}

// reference LimiterVenkatakrishnan.cpp:101-128.  The two branches there are one
// formula in the bound on the side d2 points to (max for d2 > 0, min for
// d2 < 0); selecting that bound first evaluates the same operations once per
// lane, so a wave whose cells have both signs (almost every wave) no longer
// runs the two correctly rounded divisions twice.  d2 == 0 or NaN gives 1.
__device__ __forceinline__ double face_limiter(double mn, double mx, double d2, double eps2) {
  const double m = d2 > 0 ? mx : mn;
  const double r = (1 / d2) * ((m * m + eps2) * d2 + 2.0 * d2 * d2 * m) / (m * m + 2.0 * d2 * d2 + m * d2 + eps2);
  return (d2 > 0 || d2 < 0) ? r : 1.0;
}

// the gradient, limiter and pack of one (cell, condition); the pack goes to
// stage (PW doubles) when STAGE, else to PK
template <int ND, bool STAGE>
__device__ __forceinline__ void gradlim_cell(const double *__restrict__ U, const DevMesh &dm,
                                             double *__restrict__ PK, double *__restrict__ Gout,
                                             double *__restrict__ Pout, double *__restrict__ frz,
                                             const CoefB &fmask, int i, int b, int B, double *stage) {
  constexpr int NV = ND + 2, K = ND + 1, GW = ND + 1;
  const int N = dm.N;
  int nb[K];
  double S[K][ND], w[K], u[NV], uj[K][NV];
#pragma unroll
  for (int m = 0; m < K; m++) {
    nb[m] = dm.nbr[(size_t)m * N + i];
    w[m] = dm.w[(size_t)m * N + i];
#pragma unroll
    for (int d = 0; d < ND; d++) S[m][d] = dm.Sg[((size_t)m * ND + d) * N + i];
  }
#pragma unroll
  for (int v = 0; v < NV; v++) u[v] = U[((size_t)v * N + i) * B + b];
#pragma unroll
  for (int m = 0; m < K; m++) {
    const int j = nb[m] >= 0 ? nb[m] : i;
#pragma unroll
    for (int v = 0; v < NV; v++) uj[m][v] = U[((size_t)v * N + j) * B + b];
  }
  double g[NV][ND], mn[NV], mx[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    mn[v] = u[v];
    mx[v] = u[v];
#pragma unroll
    for (int d = 0; d < ND; d++) g[v][d] = 0.0;
  }
#pragma unroll
  for (int m = 0; m < K; m++) {
    if (nb[m] >= 0) {
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const double fv = w[m] * u[v] + (1 - w[m]) * uj[m][v];
#pragma unroll
        for (int d = 0; d < ND; d++) g[v][d] += S[m][d] * fv;
        mn[v] = (uj[m][v] < mn[v]) ? uj[m][v] : mn[v];
        mx[v] = (uj[m][v] > mx[v]) ? uj[m][v] : mx[v];
      }
    } else {
#pragma unroll
      for (int v = 0; v < NV; v++)
#pragma unroll
        for (int d = 0; d < ND; d++) g[v][d] += S[m][d] * (u[v]);
      if (nb[m] == -1) {   // slip wall: tangential momentum on the face
        double nn[ND], vn = 0;
#pragma unroll
        for (int d = 0; d < ND; d++) nn[d] = dm.nf[((size_t)m * ND + d) * N + i];
#pragma unroll
        for (int d = 0; d < ND; d++) vn += u[1 + d] * nn[d];
#pragma unroll
        for (int a = 0; a < ND; a++) {
          const double tg = u[1 + a] - vn * nn[a];
#pragma unroll
          for (int d = 0; d < ND; d++) g[1 + a][d] += (tg * S[m][d] - u[1 + a] * S[m][d]);
        }
      }
    }
  }
  const double V = dm.vol[i];
  const RcpDiv divV(V);   // the gradient's nv*nd divisions by V, as one
  const double eps2 = dm.eps2[i];
  double rl[K][ND];
#pragma unroll
  for (int m = 0; m < K; m++)
#pragma unroll
    for (int d = 0; d < ND; d++) rl[m][d] = dm.rl[((size_t)m * ND + d) * N + i];
#pragma unroll
  for (int v = 0; v < NV; v++) {
#pragma unroll
    for (int d = 0; d < ND; d++) g[v][d] = divV(g[v][d]);
    double best = 1.0;
    const size_t pq = ((size_t)v * N + i) * B + b;
    if (frz && fmask.c[b] != 0.0) {
      best = frz[pq];
    } else {
      const double dmin = mn[v] - u[v], dmax = mx[v] - u[v];
#pragma unroll
      for (int m = 0; m < K; m++) {
        if (nb[m] < 0) continue;
        double d2 = 0;
#pragma unroll
        for (int d = 0; d < ND; d++) d2 += g[v][d] * rl[m][d];
        d2 /= 2.0;
        const double f = face_limiter(dmin, dmax, d2, eps2);
        best = (f < best) ? f : best;
      }
      if (frz) frz[pq] = best;
    }
    double pk[GW];
    pk[0] = u[v];
#pragma unroll
    for (int d = 0; d < ND; d++) {
      if (Gout) Gout[(((size_t)v * ND + d) * N + i) * B + b] = g[v][d];
      pk[1 + d] = best * g[v][d];
    }
    if (Pout) Pout[pq] = best;
    if constexpr (STAGE) {
#pragma unroll
      for (int q = 0; q < GW; q++) stage[v * GW + q] = pk[q];
    } else {
#pragma unroll
      for (int q = 0; q < GW; q++) PK[(((size_t)i * NV + v) * GW + q) * B + b] = pk[q];
    }
  }
}

// Green-Gauss gradient (GradientGreenGauss.cpp:21-135) fused with the
// Venkatakrishnan limiter (LimiterVenkatakrishnan.cpp:21-99).  One thread per
// (cell, condition) computes every variable; all of its loads (geometry, the
// cell's and its neighbours' states) are issued before the arithmetic that
// needs them.  Every variable's arithmetic is the reference's, in its order.
// Writes the cell pack, per variable the group [U_v | phi_v * grad U_v] at
// ((i*NV + v)*GW + q)*B + b, GW = 1 + nd.  (One wave per variable instead
// keeps fewer registers per thread but repeats the geometry loads and address
// arithmetic per variable; on the 2M-tet box it was the slower form, 427
// against 355 us, when both stored their packs per thread.)
// With frz non-null (limiter freeze on): a condition b with fmask.c[b] != 0
// takes its limiter from frz, every other condition stores its limiter there.
// With one condition the packs of the block's cells (contiguous in PK) are
// staged in LDS (blockDim.x * PW doubles of dynamic shared memory) and written
// out by the block in 16-byte pieces that cover whole cache lines; stored
// directly, each store instruction of a wave would touch 64 lines 160 bytes
// apart and lines leave L2 partly written.
// the cells of block blk (one condition), packs staged in sm (blockDim.x*PW
// doubles); every thread of the block must call it
template <int ND>
__device__ __forceinline__ void gradlim_block_one(const double *__restrict__ U, const DevMesh &dm,
                                                  double *__restrict__ PK, double *__restrict__ Gout,
                                                  double *__restrict__ Pout, double *__restrict__ frz,
                                                  const CoefB &fmask, int blk, double *sm) {
  constexpr int PW = (ND + 2) * (ND + 1);
  const int N = dm.N, i0 = blk * blockDim.x, t = i0 + threadIdx.x;
  if (t < N) gradlim_cell<ND, true>(U, dm, PK, Gout, Pout, frz, fmask, t, 0, 1, sm + threadIdx.x * PW);
  __syncthreads();
  const int nc = min((int)blockDim.x, N - i0);   // cells of this block (i0 < N: the blocks cover N)
  const double2 *src = reinterpret_cast<const double2 *>(sm);
  double2 *dst = reinterpret_cast<double2 *>(PK + (size_t)i0 * PW);
  for (int q = threadIdx.x; q < nc * (PW / 2); q += blockDim.x) dst[q] = src[q];
  __syncthreads();   // sm is free again for the block's next call
}

template <int ND, bool ONE>
__global__ __launch_bounds__(256) void k_gradlim(const double *__restrict__ U, DevMesh dm,
                                                    double *__restrict__ PK, double *__restrict__ Gout,
                                                    double *__restrict__ Pout, double *__restrict__ frz,
                                                    CoefB fmask, int lgB) {
  constexpr int NV = ND + 2, GW = ND + 1, PW = NV * GW;
  extern __shared__ double sm_pack[];
  const int N = dm.N;
  const int i0 = blockIdx.x * blockDim.x;   // dispatch order (see k_cell_residual)
  const int t = i0 + threadIdx.x;
  if constexpr (!ONE) {
    const int B = 1 << lgB, i = t >> lgB, b = t & (B - 1);
    if (i < N) gradlim_cell<ND, false>(U, dm, PK, Gout, Pout, frz, fmask, i, b, B, nullptr);
  } else {
    gradlim_block_one<ND>(U, dm, PK, Gout, Pout, frz, fmask, blockIdx.x, sm_pack);
  }
}

// a cell pack [NV][1 + ND] of one condition into registers (ONE: the
// single-condition layout, the pack in 16-byte pieces)
template <int ND, bool ONE>
__device__ __forceinline__ void load_pack(const double *__restrict__ PK, int cell, int b, int B, double *pk) {
  constexpr int PW = (ND + 2) * (ND + 1);
  if constexpr (ONE) {
    const double2 *src = reinterpret_cast<const double2 *>(PK + (size_t)cell * PW);
#pragma unroll
    for (int q = 0; q < PW / 2; q++) {
      const double2 t = src[q];
      pk[2 * q] = t.x;
      pk[2 * q + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int q = 0; q < PW; q++) pk[q] = PK[((size_t)cell * PW + q) * B + b];
  }
}

// reconstructed face state U + (phi*grad).r of one side, from its cell pack
template <int ND>
__device__ __forceinline__ void face_state(const double *pk, const double *r, double *uf) {
  constexpr int NV = ND + 2, GW = ND + 1;
#pragma unroll
  for (int v = 0; v < NV; v++) {
    uf[v] = pk[v * GW];
#pragma unroll
    for (int d = 0; d < ND; d++) uf[v] += pk[v * GW + 1 + d] * r[d];
  }
}

// One thread per (face, condition), faces in device order: second-order states
// of both sides from the cell packs, then the face flux per unit area
// (FluxRoe.tpp:306-333, FluxFreestream.tpp:192-226, FluxSlipWall.tpp:32-53),
// computed once per face and stored as a record of fs values (nv, padded to an
// even count when B = 1 so the record moves in 16-byte pieces).  finf:
// freestream primitive state per condition.
template <int ND, bool ONE>
__device__ __forceinline__ void face_flux_thread(const double *__restrict__ PK, const DevMesh &dm, const GasConst &gc,
                                                 const double *__restrict__ finf, double *__restrict__ FF, int fs,
                                                 unsigned long long *__restrict__ err, int lgB, int t) {
  constexpr int NV = ND + 2;
  const int F = dm.F, B = ONE ? 1 : 1 << lgB;
  const int f = ONE ? t : t >> lgB, b = ONE ? 0 : t & (B - 1);
  if (f >= F) return;
  const int o = dm.fo[f], j = dm.fnb[f];
  // both sides' packs and the face geometry are requested before any of them
  // is used (a boundary face reads its owner's pack twice, from cache), so the
  // two gathers overlap instead of taking one memory round trip each
  double nn[ND], rs[ND], rn[ND], ul[NV], out[NV + 1];
  double pko[(ND + 2) * (ND + 1)], pkn[(ND + 2) * (ND + 1)];
#pragma unroll
  for (int d = 0; d < ND; d++) {
    nn[d] = dm.fnr[(size_t)d * F + f];
    rs[d] = dm.frs[(size_t)d * F + f];
    rn[d] = dm.frn[(size_t)d * F + f];
  }
  load_pack<ND, ONE>(PK, o, b, B, pko);
  load_pack<ND, ONE>(PK, j >= 0 ? j : o, b, B, pkn);
  face_state<ND>(pko, rs, ul);
  if (j >= 0) {
    double ur[NV];
    face_state<ND>(pkn, rn, ur);
    roe_flux<ND, double>(gc, ul, ur, nn, out);
  } else if (j == -1) {
    slip_flux<ND, double>(gc, ul, nn, out);
  } else {
    double fi[NV + 1];
#pragma unroll
    for (int q = 0; q < NV + 1; q++) fi[q] = finf[b * (NV + 1) + q];
    freestream_flux<ND, double>(gc, ul, fi, nn, out);
  }
  bool bad = false;
#pragma unroll
  for (int v = 0; v < NV; v++) bad |= (out[v] != out[v]);
  if (bad) atomicMin(err, (unsigned long long)f);
  out[NV] = 0.0;
  if constexpr (ONE) {   // fs = NV rounded up to even
    double2 *dst = reinterpret_cast<double2 *>(FF + (size_t)f * ((NV + 1) / 2 * 2));
#pragma unroll
    for (int q = 0; q < (NV + 1) / 2; q++) dst[q] = make_double2(out[2 * q], out[2 * q + 1]);
  } else {
#pragma unroll
    for (int v = 0; v < NV; v++) FF[((size_t)f * fs + v) * B + b] = out[v];
  }
}

template <int ND, bool ONE>
__global__ __launch_bounds__(256) void k_face_flux(const double *__restrict__ PK, DevMesh dm, GasConst gc,
                                                   const double *__restrict__ finf, double *__restrict__ FF,
                                                   int fs, unsigned long long *__restrict__ err, int lgB) {
  // dispatch order (see k_cell_residual)
  face_flux_thread<ND, ONE>(PK, dm, gc, finf, FF, fs, err, lgB, blockIdx.x * blockDim.x + threadIdx.x);
}

// One thread per (cell, condition): residual sum_f +-F*A over the cell's faces
// in the reference face order (SolverExplicit.cpp:325-341), fused with the
// stage update U = U0 - alpha*dt/V*R (:364-375) and the Linf monitor (:356-362).
// The faces come pre-sorted in that order with the area negated where the cell
// is the neighbour (R + x*(-A) is R - x*A exactly).  Blocks run in dispatch
// order (consecutive blocks on different XCDs): on the 2M-tet box this kernel,
// the face flux and the gradient took 4%, 2% and 1% less time than with the
// XCD-contiguous block map of cell_index(), whose shared-L2 reuse of
// neighbour data does not pay for these gathers.
// the cells of block blk; every thread of the block must call it (the Linf
// monitor is a block reduction)
template <int ND, bool ONE>
__device__ __forceinline__ void cell_residual_block(
    // U0 and Ud may be the same array (the last stage updates in place: each
    // thread reads only its own U0 entry before writing it), so no restrict
    const double *__restrict__ FF, int fs, const double *U0, double *Ud, const double *__restrict__ dt,
    const double *__restrict__ dtg, double alpha, const DevMesh &dm, double *__restrict__ Rout,
    unsigned long long *__restrict__ linf_bits, int lgB, int blk) {
  constexpr int NV = ND + 2, K = ND + 1;
  const int N = dm.N, B = ONE ? 1 : 1 << lgB;
  const int t = blk * blockDim.x + threadIdx.x;
  const int i = ONE ? t : t >> lgB, b = ONE ? 0 : t & (B - 1);
  double R[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) R[v] = 0.0;
  if (i < N) {
    int fq[K];
    double aq[K];
#pragma unroll
    for (int q = 0; q < K; q++) {
      fq[q] = dm.cfq[(size_t)q * N + i];
      aq[q] = dm.saq[(size_t)q * N + i];
    }
#pragma unroll
    for (int q = 0; q < K; q++) {
      double x[NV + 1];
      if constexpr (ONE) {   // fs = NV rounded up to even
        const double2 *src = reinterpret_cast<const double2 *>(FF + (size_t)fq[q] * ((NV + 1) / 2 * 2));
#pragma unroll
        for (int p = 0; p < (NV + 1) / 2; p++) {
          const double2 tt = src[p];
          x[2 * p] = tt.x;
          x[2 * p + 1] = tt.y;
        }
      } else {
#pragma unroll
        for (int v = 0; v < NV; v++) x[v] = FF[((size_t)fq[q] * fs + v) * B + b];
      }
#pragma unroll
      for (int v = 0; v < NV; v++) R[v] += x[v] * aq[q];
    }
    if (Ud) {
      const double c = alpha * (dtg ? dtg[b] : dt[(size_t)i * B + b]) / dm.vol[i];
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const size_t q = ((size_t)v * N + i) * B + b;
        Ud[q] = U0[q] - c * R[v];
      }
    }
    if (Rout) {
#pragma unroll
      for (int v = 0; v < NV; v++) Rout[((size_t)v * N + i) * B + b] = R[v];
    }
  }
  if (linf_bits) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      double a = block_max_cond(::fabs(R[v]), B);
      if ((int)threadIdx.x < B)
        atomicMax(&linf_slot_of(linf_bits, NV * B, blk)[v * B + threadIdx.x],
                  (unsigned long long)__double_as_longlong(a));
    }
  }
}

template <int ND, bool ONE>
__global__ __launch_bounds__(256) void k_cell_residual(const double *__restrict__ FF, int fs, const double *U0,
                                                       double *Ud, const double *__restrict__ dt,
                                                       const double *__restrict__ dtg, double alpha, DevMesh dm,
                                                       double *__restrict__ Rout,
                                                       unsigned long long *__restrict__ linf_bits, int lgB) {
  cell_residual_block<ND, ONE>(FF, fs, U0, Ud, dt, dtg, alpha, dm, Rout, linf_bits, lgB, blockIdx.x);
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Linf bit patterns [v][B] -> out[(it*B + b)*nv + v]; the row index comes from
// the device counter *dit when given (graph replays), which is then advanced
__global__ void k_linf_decode(const unsigned long long *bits, double *out, int nv, int B, int it, int cap,
                              int *dit) {
  const int q = threadIdx.x;
  if (dit) it = *dit;
  __syncthreads();
  if (it < cap && q < nv * B) {
    const int v = q / B, b = q - v * B;
    out[((size_t)it * B + b) * nv + v] = __longlong_as_double((long long)linf_slots_max(bits, nv * B, q));
  }
  if (dit && q == 0) *dit = it + 1;
}

// bits[q] = max over the Linf copies (copy 0 in place)
__global__ void k_linf_reduce(unsigned long long *bits, int nq) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nq) {
    const unsigned long long m = linf_slots_max(bits, nq, q);
    bits[q] = m;
  }
}

// Persistent form of bes_explicit_run for one condition with local time
// steps: one cooperative launch (every block resident) runs all niter
// pseudo-time steps, the passes of each (time step; per stage gradient, face
// fluxes, residual and update) separated by grid barriers, each pass a loop
// of the grid's blocks over the pass's blocks of 256 cells or faces.  Every
// value is computed by the same device code as the per-pass kernels, so the
// states and monitors are bitwise those of the launched form.  A grid barrier
// replaces each launch (14 per RK4 step); on one MI355X the barrier costs more
// than a graph-launched kernel on a 40k-cell mesh, so the graph stays the
// default and this form is an option.
struct PersistStep {
  double *U0, *Ua, *Ub;   // state slots as in explicit_iteration
  double *PK, *FF, *dt, *frz, *linf;
  const double *finf;
  unsigned long long *lbits, *err;
  double cfl, alpha[16];
  int nstage, niter, linf_cap;
};

template <int ND>
__global__ __launch_bounds__(256) void k_explicit_persistent(PersistStep a, DevMesh dm, GasConst gc, CoefB fmask) {
  constexpr int NV = ND + 2, FS = (NV + 1) / 2 * 2;
  extern __shared__ double sm_pack[];
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const int N = dm.N, nbc = (N + blockDim.x - 1) / blockDim.x, nbf = (dm.F + blockDim.x - 1) / blockDim.x;
  double *U0 = a.U0, *Ua = a.Ua;
  for (int it = 0; it < a.niter; it++) {
    for (int blk = blockIdx.x; blk < nbc; blk += gridDim.x) {
      const int i = blk * blockDim.x + threadIdx.x;
      if (i < N) timestep_cell<ND>(U0, dm, gc, a.cfl, a.dt, i, 0, 1);
    }
    grid.sync();
    for (int st = 0; st < a.nstage; st++) {
      const bool last = st == a.nstage - 1;
      const double *src = st == 0 ? U0 : ((st - 1) % 2 == 0 ? Ua : a.Ub);
      double *dst = (last && a.nstage > 1) ? U0 : (st % 2 == 0 ? Ua : a.Ub);
      for (int blk = blockIdx.x; blk < nbc; blk += gridDim.x)
        gradlim_block_one<ND>(src, dm, a.PK, nullptr, nullptr, a.frz, fmask, blk, sm_pack);
      grid.sync();
      // the monitor words are cleared here: block 0 decoded the previous
      // step's before it reached the barrier above, and the adds come after
      // the next one
      if (st == 0 && a.linf)
        for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < kLinfSlots * NV; q += gridDim.x * blockDim.x)
          a.lbits[q] = 0ull;
      for (int blk = blockIdx.x; blk < nbf; blk += gridDim.x)
        face_flux_thread<ND, true>(a.PK, dm, gc, a.finf, a.FF, FS, a.err, 0, blk * blockDim.x + threadIdx.x);
      grid.sync();
      for (int blk = blockIdx.x; blk < nbc; blk += gridDim.x)
        cell_residual_block<ND, true>(a.FF, FS, U0, dst, a.dt, nullptr, a.alpha[st], dm, nullptr,
                                      (last && a.linf) ? a.lbits : nullptr, 0, blk);
      grid.sync();
    }
    if (a.linf && blockIdx.x == 0 && (int)threadIdx.x < NV && it < a.linf_cap)
      a.linf[(size_t)it * NV + threadIdx.x] =
          __longlong_as_double((long long)linf_slots_max(a.lbits, NV, threadIdx.x));
    if (a.nstage == 1) {
      double *t = U0;
      U0 = Ua;
      Ua = t;
    }
  }
}

// per-face flux / Jacobian batches (host-facing known-answer interface)
template <int ND>
__global__ void k_flux_batch(int kind, int n, const double *__restrict__ UL,
                             const double *__restrict__ UR, const double *__restrict__ nrm,
                             GasConst gc, double *__restrict__ F) {
  constexpr int NV = ND + 2;
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  double ul[NV], ur[NV + 1], nn[ND], out[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) ul[v] = UL[(size_t)f * NV + v];
#pragma unroll
  for (int v = 0; v < NV + 1; v++) ur[v] = UR[(size_t)f * (NV + 1) + v];
#pragma unroll
  for (int d = 0; d < ND; d++) nn[d] = nrm[(size_t)f * ND + d];
  if (kind == kInternal) roe_flux<ND, double>(gc, ul, ur, nn, out);
  else if (kind == kSlipWall) slip_flux<ND, double>(gc, ul, nn, out);
  else freestream_flux<ND, double>(gc, ul, ur, nn, out);
#pragma unroll
  for (int v = 0; v < NV; v++) F[(size_t)f * NV + v] = out[v];
}

// one thread per (face, input direction); 2*NV directions for Roe, NV otherwise
template <int ND>
__global__ void k_jacobian_batch(int kind, int n, const double *__restrict__ UL,
                                 const double *__restrict__ UR, const double *__restrict__ nrm,
                                 GasConst gc, double *__restrict__ JL, double *__restrict__ JR) {
  constexpr int NV = ND + 2;
  const int ndir = 2 * NV;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int f = t / ndir, dir = t % ndir;
  if (f >= n) return;
  if (kind != kInternal && dir >= NV) return;
  Dual ul[NV], ur[NV + 1], out[NV];
  double nn[ND];
#pragma unroll
  for (int v = 0; v < NV; v++) ul[v] = Dual(UL[(size_t)f * NV + v], dir == v ? 1.0 : 0.0);
#pragma unroll
  for (int v = 0; v < NV + 1; v++) ur[v] = Dual(UR[(size_t)f * (NV + 1) + v], dir == NV + v ? 1.0 : 0.0);
#pragma unroll
  for (int d = 0; d < ND; d++) nn[d] = nrm[(size_t)f * ND + d];
  if (kind == kInternal) roe_flux<ND, Dual>(gc, ul, ur, nn, out);
  else if (kind == kSlipWall) slip_flux<ND, Dual>(gc, ul, nn, out);
  else freestream_flux<ND, Dual>(gc, ul, ur, nn, out);
  double *J = dir < NV ? JL : JR;
  int col = dir < NV ? dir : dir - NV;
  if (!J) return;
#pragma unroll
  for (int o = 0; o < NV; o++) J[((size_t)f * NV + o) * NV + col] = out[o].d;
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

template <typename T>
static int dalloc(T **p, size_t n) {
  BES_HIP(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
  return BES_OK;
}

template <typename T>
static int upload(T **p, const std::vector<T> &h) {
  int s = dalloc(p, h.size());
  if (s) return s;
  if (!h.empty()) BES_HIP(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return BES_OK;
}

static GasConst make_gas(const bes_gas *gas) {
  GasConst g;
  g.gamma = gas ? gas->gamma : 1.4;
  g.R = gas ? gas->R : 287.0;
  g.harten = gas ? gas->harten : 0.05;
  g.cv = g.R / (g.gamma - 1.0);
  g.cp = g.cv + g.R;
  return g;
}

// Device layout from the reference geometry (cells in Morton order).
static int build_device_mesh(bes_ctx *c) {
  const Geometry &g = c->geo;
  const int nd = g.nd, k = g.k, N = (int)g.ncell;
  c->perm = locality_order(g);
  c->iperm.assign(N, 0);
  for (int i = 0; i < N; i++) c->iperm[c->perm[i]] = i;
  std::vector<double> vol(N), eps2(N), sproj((size_t)nd * N), Sg((size_t)k * nd * N), w((size_t)k * N),
      rl((size_t)k * nd * N), nf((size_t)k * nd * N), af((size_t)k * N);
  std::vector<int> nbr((size_t)k * N), cfd((size_t)k * N), cfq((size_t)k * N);
  std::vector<double> saq((size_t)k * N);
  std::vector<uint32_t> info(N);
  // faces in device order: by owner (Morton index), then the owner's local face
  const int F = (int)g.nface;
  std::vector<int64_t> dev2ref;
  dev2ref.reserve(F);
  std::vector<int> ref2dev(F, -1);
  for (int i = 0; i < N; i++) {
    const int64_t e = c->perm[i];
    for (int m = 0; m < k; m++) {
      const int64_t f = g.cface[e * k + m];
      if (g.owner[f] == e && ref2dev[f] < 0) {
        ref2dev[f] = (int)dev2ref.size();
        dev2ref.push_back(f);
      }
    }
  }
  if ((int)dev2ref.size() != F) { set_error("face ownership is inconsistent"); return BES_ERR_MESH; }
  std::vector<int> fo(F), fnb(F);
  std::vector<double> fnr((size_t)nd * F), frs((size_t)nd * F), frn((size_t)nd * F, 0.0);
  // reconstruction vector a cell uses for one of its faces (rsel: Q22 in compat mode)
  auto recon = [&](int64_t cell, int64_t f, int d) -> double {
    for (int mm = 0; mm < k; mm++)
      if (g.cface[cell * k + mm] == f) return g.r[(cell * k + g.rsel[cell * k + mm]) * nd + d];
    return 0.0;
  };
  for (int q = 0; q < F; q++) {
    const int64_t f = dev2ref[q];
    fo[q] = (int)c->iperm[g.owner[f]];
    fnb[q] = g.nbr[f] >= 0 ? (int)c->iperm[g.nbr[f]] : (g.btype[f] == kSlipWall ? -1 : -2);
    for (int d = 0; d < nd; d++) {
      fnr[(size_t)d * F + q] = g.n[f * nd + d];
      frs[(size_t)d * F + q] = recon(g.owner[f], f, d);
      if (g.nbr[f] >= 0) frn[(size_t)d * F + q] = recon(g.nbr[f], f, d);
    }
  }
  c->fref = dev2ref;
  for (int i = 0; i < N; i++) {
    const int64_t e = c->perm[i];
    vol[i] = g.vol[e];
    // LimiterVenkatakrishnan.cpp:42-46, epsilon squared where it is used
    double L = std::pow(g.vol[e], 1.0 / nd);
    double eps = std::pow(c->venkat_k * L, (3.0 / 2.0));
    eps2[i] = eps * eps;
    for (int d = 0; d < nd; d++) sproj[(size_t)d * N + i] = g.sproj[e * nd + d];
    int order[4];
    for (int m = 0; m < k; m++) order[m] = m;
    std::sort(order, order + k, [&](int a, int b) { return g.cface[e * k + a] < g.cface[e * k + b]; });
    uint32_t bits = 0;
    for (int q = 0; q < k; q++) bits |= (uint32_t)order[q] << (2 * q);
    for (int q = 0; q < k; q++) {   // faces in reference order, signed area
      const int64_t f = g.cface[e * k + order[q]];
      cfq[(size_t)q * N + i] = ref2dev[f];
      saq[(size_t)q * N + i] = g.owner[f] == e ? g.area[f] : -g.area[f];
    }
    for (int m = 0; m < k; m++) {
      const int64_t s = e * k + m, f = g.cface[s];
      const bool own = g.owner[f] == e;
      if (own) bits |= 1u << (8 + m);
      const int64_t jadj = g.cadj[s];
      int code;
      if (jadj >= 0) code = (int)c->iperm[jadj];
      else code = (g.btype[f] == kSlipWall) ? -1 : -2;
      nbr[(size_t)m * N + i] = code;
      w[(size_t)m * N + i] = g.w[s];
      af[(size_t)m * N + i] = g.area[f];
      cfd[(size_t)m * N + i] = ref2dev[f];
      for (int d = 0; d < nd; d++) {
        const size_t q = ((size_t)m * nd + d) * N + i;
        Sg[q] = own ? g.S[f * nd + d] : -g.S[f * nd + d];
        rl[q] = g.r[s * nd + d];
        nf[q] = g.n[f * nd + d];
      }
    }
    info[i] = bits;
  }
  DevMesh &dm = c->dm;
  dm.nd = nd; dm.k = k; dm.N = N; dm.F = F;
  int s;
  if ((s = upload(&dm.vol, vol)) || (s = upload(&dm.eps2, eps2)) || (s = upload(&dm.sproj, sproj)) ||
      (s = upload(&dm.nbr, nbr)) || (s = upload(&dm.Sg, Sg)) || (s = upload(&dm.w, w)) ||
      (s = upload(&dm.rl, rl)) || (s = upload(&dm.nf, nf)) || (s = upload(&dm.af, af)) ||
      (s = upload(&dm.cfd, cfd)) || (s = upload(&dm.info, info)) || (s = upload(&dm.cfq, cfq)) ||
      (s = upload(&dm.saq, saq)) || (s = upload(&dm.fo, fo)) ||
      (s = upload(&dm.fnb, fnb)) || (s = upload(&dm.fnr, fnr)) || (s = upload(&dm.frs, frs)) ||
      (s = upload(&dm.frn, frn)))
    return s;
  return BES_OK;
}

static int upload_finf(bes_ctx *c) {
  BES_HIP(hipMemcpyAsync(c->dfinf, c->finf_b.data(), c->finf_b.size() * sizeof(double), hipMemcpyHostToDevice,
                         c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  return BES_OK;
}

static int create_from_input(const MeshInput &in, int geom_mode, const bes_gas *gas, bes_ctx **out) {
  if (!out) { set_error("null output pointer"); return BES_ERR_ARG; }
  *out = nullptr;
  if (geom_mode < 0 || geom_mode > 7) { set_error("geom_mode must be in [0,7]"); return BES_ERR_ARG; }
  bes_ctx *c = new bes_ctx();
  try {
    c->geo = build_geometry(in, geom_mode);
  } catch (const std::exception &ex) {
    set_error(ex.what());
    delete c;
    return BES_ERR_MESH;
  }
  if (c->geo.ncell > 0x7fffffff / 4) { set_error("mesh too large"); delete c; return BES_ERR_ARG; }
  c->nd = c->geo.nd; c->nv = c->nd + 2; c->k = c->geo.k; c->N = (int)c->geo.ncell;
  c->node_xyz = in.coords;
  c->cell_nodes = in.cells;
  c->gc = make_gas(gas);
  c->venkat_k = gas ? gas->venkat_k : 5.0;
  int s;
  hipError_t he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (he != hipSuccess) { s = hip_fail(he, "hipStreamCreate"); delete c; return s; }
  const size_t N = c->N, nv = c->nv, nd = c->nd;
  c->fs = (int)((nv + 1) / 2 * 2);
  if ((s = build_device_mesh(c)) || (s = dalloc(&c->U0, nv * N)) || (s = dalloc(&c->Ua, nv * N)) ||
      (s = dalloc(&c->Ub, nv * N)) || (s = dalloc(&c->LG, nv * (1 + nd) * N)) ||
      (s = dalloc(&c->FF, (size_t)c->fs * c->dm.F)) || (s = dalloc(&c->dt, N)) ||
      (s = dalloc(&c->dtg, bes_ctx::kMaxBatch)) || (s = dalloc(&c->red, bes_ctx::kRedWords)) ||
      (s = dalloc(&c->lbits, (size_t)kLinfSlots * 5 * bes_ctx::kMaxBatch)) ||
      (s = dalloc(&c->err, 1)) || (s = dalloc(&c->dfinf, bes_ctx::kMaxBatch * 6)) ||
      (s = dalloc(&c->dit, 1))) {
    bes_ctx_destroy(c);
    return s;
  }
  // default freestream: Mach 0 (Config.cpp:481-487 form)
  double fi[6] = {1.0, 0, 0, 0, 0, 0};
  fi[nd + 1] = 1.0 / c->gc.gamma;
  fi[nd + 2] = 1.0 / (c->gc.gamma * c->gc.R);
  std::memcpy(c->finf, fi, sizeof(fi));
  c->finf_b.assign(fi, fi + nv + 1);
  if ((s = upload_finf(c))) { bes_ctx_destroy(c); return s; }
  *out = c;
  return BES_OK;
}

static int reset_err(bes_ctx *c) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, c->err, 1, ~0ull);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

static int check_err(bes_ctx *c) {
  unsigned long long e = 0;
  BES_HIP(hipMemcpyAsync(&e, c->err, sizeof(e), hipMemcpyDeviceToHost, c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  if (e != ~0ull) {
    c->last_err_face = e < c->fref.size() ? (long)c->fref[e] : -1;   // device face -> reference id
    set_error("non-finite flux at face " + std::to_string(c->last_err_face));
    return BES_ERR_NAN;
  }
  return BES_OK;
}

// launch helpers shared with the implicit path
int launch_timestep(bes_ctx *c, const double *U, double cfl, int local, const double *cfl_b = nullptr) {
  const int bs = tuned_bs(c, kProfTimestep), B = c->B, nb = grid_for(c->N * B, bs);
  CoefB cb;
  for (int b = 0; b < B; b++) cb.c[b] = cfl_b ? cfl_b[b] : cfl;
  const int pc = cfl_b ? 1 : 0;
  unsigned long long *gm = nullptr;
  if (!local) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, c->red, B, 0x7ff0000000000000ull);
    gm = c->red;
  }
  {
    ProfScope ps(c, kProfTimestep, (double)c->N * (8.0 * (c->nv * B + c->nd + 1) + 8.0 * B));
    if (c->nd == 2)
      hipLaunchKernelGGL(k_timestep<2>, dim3(nb), dim3(bs), 0, c->stream, U, c->dm, c->gc, cfl, cb, pc, c->dt, gm,
                         c->lgB);
    else
      hipLaunchKernelGGL(k_timestep<3>, dim3(nb), dim3(bs), 0, c->stream, U, c->dm, c->gc, cfl, cb, pc, c->dt, gm,
                         c->lgB);
  }
  if (!local) hipLaunchKernelGGL(k_decode_min, dim3(1), dim3(64), 0, c->stream, c->red, c->dtg, B);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

// calls f(integral_constant nd, bool_constant one) for the kernel instance
// of the mesh dimension and of a single-condition (one) or batched launch
template <typename Fn>
void by_shape(int nd, bool one, Fn &&f) {
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  if (nd == 2) {
    if (one) f(I2{}, std::true_type{});
    else f(I2{}, std::false_type{});
  } else {
    if (one) f(I3{}, std::true_type{});
    else f(I3{}, std::false_type{});
  }
}

// algorithmic HBM bytes of one gradient+limiter launch: state, geometry, pack
double gradlim_bytes(const bes_ctx *c) {
  const double N = c->N, B = c->B, nv = c->nv, nd = c->nd, k = c->k;
  return N * (8 * nv * B + k * (4 + 16 * nd + 8) + 16 + 8 * nv * (1 + nd) * B);
}

// one thread per (cell, condition)
int launch_gradlim(bes_ctx *c, const double *U, double *Gout, double *Pout) {
  const int bs = tuned_bs(c, kProfGradlim), nb = grid_for(c->N * c->B, bs);
  ProfScope ps(c, kProfGradlim, gradlim_bytes(c));
  CoefB fm;
  for (int b = 0; b < bes_ctx::kMaxBatch; b++)
    fm.c[b] = (b < (int)c->lim_frozen.size() && c->lim_frozen[b]) ? 1.0 : 0.0;
  double *frz = c->lim_freeze_on ? c->phi_frz : nullptr;
  // one condition: the block's packs are staged in LDS
  const bool one = c->B == 1;
  const size_t lds = one ? (size_t)bs * c->nv * (c->nd + 1) * sizeof(double) : 0;
  by_shape(c->nd, one, [&](auto nd, auto one_) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gradlim<decltype(nd)::value, decltype(one_)::value>),
                       dim3(nb), dim3(bs), lds, c->stream, U, c->dm, c->LG, Gout, Pout, frz, fm, c->lgB);
  });
  BES_HIP(hipGetLastError());
  return BES_OK;
}

// face flux: face records, both cell packs (each pack read once overall), fluxes
double faceflux_bytes(const bes_ctx *c) {
  const double F = c->dm.F, N = c->N, B = c->B, nd = c->nd, nv = c->nv;
  return F * (8 + 24 * nd + 8 * c->fs * B) + N * 8 * nv * (1 + nd) * B;
}

// face fluxes of the current packs (shared with the implicit assembly)
int launch_face_flux(bes_ctx *c) {
  const int bs = tuned_bs(c, kProfFaceFlux), nb = grid_for(c->dm.F * c->B, bs);
  const bool one = c->B == 1;
  ProfScope ps(c, kProfFaceFlux, faceflux_bytes(c));
  by_shape(c->nd, one, [&](auto nd, auto one_) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_face_flux<decltype(nd)::value, decltype(one_)::value>),
                       dim3(nb), dim3(bs), 0, c->stream, c->LG, c->dm, c->gc, c->dfinf, c->FF, c->fs, c->err,
                       c->lgB);
  });
  BES_HIP(hipGetLastError());
  return BES_OK;
}

// per cell: ordered face ids and signed areas, dt, volume, U0 and U; the face
// flux records once
double residual_bytes(const bes_ctx *c, bool update, bool rout) {
  const double N = c->N, B = c->B, nv = c->nv, k = c->k;
  return N * (k * 12 + (update ? 8 + 8 * B + 16 * nv * B : 0) + (rout ? 8 * nv * B : 0)) +
         (double)c->dm.F * 8 * c->fs * B;
}

// residual of the packs of the last gradient pass; with U0/Ud the stage update
int launch_residual(bes_ctx *c, const double *U0, double *Ud, double alpha, int local, double *Rout,
                    unsigned long long *linf_bits) {
  int s = launch_face_flux(c);
  if (s) return s;
  const int bs = tuned_bs(c, kProfResidual), nb = grid_for(c->N * c->B, bs);
  const double *dtg = local ? nullptr : c->dtg;
  const bool one = c->B == 1;
  ProfScope ps(c, kProfResidual, residual_bytes(c, U0 != nullptr, Rout != nullptr));
  by_shape(c->nd, one, [&](auto nd, auto one_) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_residual<decltype(nd)::value, decltype(one_)::value>),
                       dim3(nb), dim3(bs), 0, c->stream, c->FF, c->fs, U0, Ud, c->dt, dtg, alpha, c->dm, Rout,
                       linf_bits, c->lgB);
  });
  BES_HIP(hipGetLastError());
  return BES_OK;
}

// device (Morton order, [comp][N][B]) <-> host (reference order, [B][ncell][comp])
int state_to_host(bes_ctx *c, const double *dsoa, int ncomp, double *host) {
  const size_t N = c->N, B = c->B;
  std::vector<double> tmp(ncomp * N * B);
  BES_HIP(hipMemcpyAsync(tmp.data(), dsoa, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  for (size_t b = 0; b < B; b++)
    for (size_t i = 0; i < N; i++) {
      const int64_t e = c->perm[i];
      for (int q = 0; q < ncomp; q++) host[(b * N + e) * ncomp + q] = tmp[((size_t)q * N + i) * B + b];
    }
  return BES_OK;
}

int state_to_device(bes_ctx *c, const double *host, int ncomp, double *dsoa) {
  const size_t N = c->N, B = c->B;
  std::vector<double> tmp(ncomp * N * B);
  for (size_t b = 0; b < B; b++)
    for (size_t i = 0; i < N; i++) {
      const int64_t e = c->perm[i];
      for (int q = 0; q < ncomp; q++) tmp[((size_t)q * N + i) * B + b] = host[(b * N + e) * ncomp + q];
    }
  BES_HIP(hipMemcpyAsync(dsoa, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  return BES_OK;
}

static void drop_graph(bes_ctx *c) {
  if (c->graph) hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
}

// (Re)allocate every per-condition array for a batch of B conditions
static int resize_batch(bes_ctx *c, int B) {
  if (c->stream) BES_HIP(hipStreamSynchronize(c->stream));
  drop_graph(c);
  double **arrs[] = {&c->U0, &c->Ua, &c->Ub, &c->LG, &c->FF, &c->dt, &c->G, &c->phi, &c->R, &c->linf, &c->phi_frz};
  for (double **p : arrs) {
    if (*p) hipFree(*p);
    *p = nullptr;
  }
  c->lim_freeze_drop = 0.0;   // the freeze is per batch: set it again after a batch change
  c->lim_freeze_on = false;
  c->lim_frozen.clear();
  c->lim_epoch++;
  c->linf_cap = 0;
  c->B = B;
  c->lgB = 0;
  while ((1 << c->lgB) < B) c->lgB++;
  const size_t N = c->N, nv = c->nv, nd = c->nd, Bs = B;
  int s;
  c->fs = B == 1 ? (int)((nv + 1) / 2 * 2) : (int)nv;
  if ((s = dalloc(&c->U0, nv * N * Bs)) || (s = dalloc(&c->Ua, nv * N * Bs)) || (s = dalloc(&c->Ub, nv * N * Bs)) ||
      (s = dalloc(&c->LG, nv * (1 + nd) * N * Bs)) || (s = dalloc(&c->FF, (size_t)c->fs * c->dm.F * Bs)) ||
      (s = dalloc(&c->dt, N * Bs)))
    return s;
  std::vector<double> fb((size_t)B * (nv + 1));
  for (int b = 0; b < B; b++) std::copy(c->finf, c->finf + nv + 1, fb.begin() + (size_t)b * (nv + 1));
  c->finf_b = fb;
  return upload_finf(c);
}

}  // namespace bes

using namespace bes;

// ---------------------------------------------------------------------------
// C ABI (declared in include/bes_hip.h)
// ---------------------------------------------------------------------------
extern "C" {

const char *bes_last_error_message(void) { return g_err.c_str(); }

const char *bes_status_string(int s) {
  switch (s) {
    case BES_OK: return "ok";
    case BES_ERR_ARG: return "invalid argument";
    case BES_ERR_HIP: return "HIP runtime error";
    case BES_ERR_NAN: return "non-finite flux";
    case BES_ERR_MESH: return "mesh error";
    case BES_ERR_SOLVE: return "linear solver did not converge";
    default: return "unknown status";
  }
}

int bes_ctx_create(const bes_mesh_desc *mesh, int geom_mode, const bes_gas *gas, bes_ctx **out) {
  if (!mesh || !mesh->coords || !mesh->cells || (mesh->nbface > 0 && (!mesh->bfaces || !mesh->btype)) ||
      (mesh->ndim != 2 && mesh->ndim != 3) || mesh->ncell <= 0 || mesh->nnode <= 0) {
    set_error("invalid mesh descriptor");
    return BES_ERR_ARG;
  }
  MeshInput in;
  in.ndim = mesh->ndim;
  in.coords.assign(mesh->coords, mesh->coords + mesh->nnode * mesh->ndim);
  in.cells.assign(mesh->cells, mesh->cells + mesh->ncell * (mesh->ndim + 1));
  if (mesh->nbface > 0) {
    in.bfaces.assign(mesh->bfaces, mesh->bfaces + mesh->nbface * mesh->ndim);
    in.btype.assign(mesh->btype, mesh->btype + mesh->nbface);
  }
  return create_from_input(in, geom_mode, gas, out);
}

int bes_ctx_create_msh2(const char *path, int geom_mode, const bes_gas *gas, bes_ctx **out) {
  if (!path) { set_error("null path"); return BES_ERR_ARG; }
  MeshInput in;
  try {
    in = read_msh2(path);
  } catch (const std::exception &ex) {
    set_error(ex.what());
    return BES_ERR_MESH;
  }
  return create_from_input(in, geom_mode, gas, out);
}

void bes_implicit_free(bes_ctx *c);   // implicit.hip

int bes_ctx_destroy(bes_ctx *c) {
  if (!c) return BES_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  bes_implicit_free(c);
  void *ptrs[] = {c->dm.vol, c->dm.eps2, c->dm.sproj, c->dm.nbr, c->dm.Sg, c->dm.w, c->dm.rl, c->dm.nf,
                  c->dm.af, c->dm.cfd, c->dm.info, c->dm.cfq, c->dm.saq, c->dm.fo, c->dm.fnb, c->dm.fnr,
                  c->dm.frs, c->dm.frn,
                  c->FF, c->U0, c->Ua, c->Ub, c->LG, c->G, c->phi, c->phi_frz,
                  c->dt, c->dtg, c->R, c->red, c->lbits, c->err, c->linf, c->dfinf, c->dit};
  drop_graph(c);
  for (void *p : ptrs)
    if (p) hipFree(p);
  for (auto &p : c->ev_pending) { c->ev_pool.push_back(p.a); c->ev_pool.push_back(p.b); }
  for (hipEvent_t e : c->ev_pool) if (e) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return BES_OK;
}

int bes_ctx_info(const bes_ctx *c, int *ndim, long *ncell, long *nface) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  if (ndim) *ndim = c->nd;
  if (ncell) *ncell = (long)c->geo.ncell;
  if (nface) *nface = (long)c->geo.nface;
  return BES_OK;
}

int bes_get_geometry(const bes_ctx *c, long *owner, long *nbr, int *btype, double *S, double *n,
                     double *area, double *vol, double *centroid, long *cell_faces, long *cell_adj,
                     double *w, double *r) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  const Geometry &g = c->geo;
  if (owner) std::copy(g.owner.begin(), g.owner.end(), owner);
  if (nbr) std::copy(g.nbr.begin(), g.nbr.end(), nbr);
  if (btype) std::copy(g.btype.begin(), g.btype.end(), btype);
  if (S) std::copy(g.S.begin(), g.S.end(), S);
  if (n) std::copy(g.n.begin(), g.n.end(), n);
  if (area) std::copy(g.area.begin(), g.area.end(), area);
  if (vol) std::copy(g.vol.begin(), g.vol.end(), vol);
  if (centroid) std::copy(g.centroid.begin(), g.centroid.end(), centroid);
  if (cell_faces) std::copy(g.cface.begin(), g.cface.end(), cell_faces);
  if (cell_adj) std::copy(g.cadj.begin(), g.cadj.end(), cell_adj);
  if (w) std::copy(g.w.begin(), g.w.end(), w);
  if (r) std::copy(g.r.begin(), g.r.end(), r);
  return BES_OK;
}

int bes_set_freestream(bes_ctx *c, const double *prim) {
  if (!c || !prim) { set_error("null argument"); return BES_ERR_ARG; }
  const int nv = c->nv;
  for (int q = 0; q < nv + 1; q++) c->finf[q] = prim[q];
  for (int b = 0; b < c->B; b++) std::copy(prim, prim + nv + 1, c->finf_b.begin() + (size_t)b * (nv + 1));
  return upload_finf(c);
}

int bes_set_batch(bes_ctx *c, int nbatch) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  if (nbatch < 1 || nbatch > bes_ctx::kMaxBatch || (nbatch & (nbatch - 1))) {
    set_error("batch size must be a power of two in [1, 64]");
    return BES_ERR_ARG;
  }
  if ((long)c->dm.F * nbatch > 0x7fffffffL || (long)c->N * nbatch > 0x7fffffffL) {
    set_error("mesh too large for this batch size");
    return BES_ERR_ARG;
  }
  if (nbatch == c->B) return BES_OK;
  bes_implicit_free(c);   // the implicit workspace is sized for one batch size
  return resize_batch(c, nbatch);
}

int bes_set_limiter_freeze(bes_ctx *c, double drop) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  if (!(drop >= 0.0 && drop < 1.0)) { set_error("limiter freeze drop must be in [0, 1)"); return BES_ERR_ARG; }
  BES_HIP(hipStreamSynchronize(c->stream));
  if (drop > 0.0 && !c->phi_frz) {
    int s = dalloc(&c->phi_frz, (size_t)c->nv * c->N * c->B);
    if (s) return s;
  }
  c->lim_freeze_drop = drop;
  c->lim_freeze_on = drop > 0.0;
  c->lim_frozen.assign(c->B, 0);
  c->lim_epoch++;
  return BES_OK;
}

int bes_freeze_limiter_values(bes_ctx *c, const double *phi) {
  if (!c || !phi) { set_error("null argument"); return BES_ERR_ARG; }
  BES_HIP(hipStreamSynchronize(c->stream));
  if (!c->phi_frz) {
    int s = dalloc(&c->phi_frz, (size_t)c->nv * c->N * c->B);
    if (s) return s;
  }
  int s = state_to_device(c, phi, c->nv, c->phi_frz);
  if (s) return s;
  c->lim_freeze_on = true;
  c->lim_freeze_drop = 0.0;
  c->lim_frozen.assign(c->B, 1);
  c->lim_epoch++;
  return BES_OK;
}

int bes_limiter_frozen(bes_ctx *c, int *frozen) {
  if (!c || !frozen) { set_error("invalid argument"); return BES_ERR_ARG; }
  for (int b = 0; b < c->B; b++) frozen[b] = b < (int)c->lim_frozen.size() && c->lim_frozen[b];
  return BES_OK;
}

int bes_get_batch(const bes_ctx *c) { return c ? c->B : 0; }

int bes_set_freestream_batch(bes_ctx *c, const double *prim) {
  if (!c || !prim) { set_error("null argument"); return BES_ERR_ARG; }
  const int nv = c->nv;
  std::copy(prim, prim + (size_t)c->B * (nv + 1), c->finf_b.begin());
  for (int q = 0; q < nv + 1; q++) c->finf[q] = prim[q];
  return upload_finf(c);
}

int bes_set_state(bes_ctx *c, const double *U) {
  if (!c || !U) { set_error("null argument"); return BES_ERR_ARG; }
  return state_to_device(c, U, c->nv, c->U0);
}

int bes_get_state(bes_ctx *c, double *U) {
  if (!c || !U) { set_error("null argument"); return BES_ERR_ARG; }
  return state_to_host(c, c->U0, c->nv, U);
}

int bes_compute_time_step(bes_ctx *c, double cfl, int local) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  int s = launch_timestep(c, c->U0, cfl, local);
  if (s) return s;
  c->dt_local = local != 0;
  BES_HIP(hipStreamSynchronize(c->stream));
  return BES_OK;
}

// TimeStepCalculator::getTimeStep for every element (the global minimum on
// every element after a global computation, TimeStepCalculatorGlobal.cpp:100-102)
int bes_get_time_step(bes_ctx *c, double *dt) {
  if (!c || !dt) { set_error("null argument"); return BES_ERR_ARG; }
  if (c->dt_local) return state_to_host(c, c->dt, 1, dt);
  std::vector<double> g(c->B);
  BES_HIP(hipMemcpy(g.data(), c->dtg, c->B * sizeof(double), hipMemcpyDeviceToHost));
  for (int b = 0; b < c->B; b++) std::fill(dt + (size_t)b * c->N, dt + (size_t)(b + 1) * c->N, g[b]);
  return BES_OK;
}

int bes_compute_gradient_limiter(bes_ctx *c) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  const size_t N = c->N, B = c->B;
  if (!c->G) {
    int s;
    if ((s = dalloc(&c->G, (size_t)c->nv * c->nd * N * B)) || (s = dalloc(&c->phi, (size_t)c->nv * N * B))) return s;
  }
  int s = launch_gradlim(c, c->U0, c->G, c->phi);
  if (s) return s;
  BES_HIP(hipStreamSynchronize(c->stream));
  return BES_OK;
}

int bes_get_gradient(bes_ctx *c, double *G) {
  if (!c || !G || !c->G) { set_error("gradient not computed"); return BES_ERR_ARG; }
  // device [v][d][N][B] -> host [B][cell][v][d]
  return state_to_host(c, c->G, c->nv * c->nd, G);
}

int bes_get_limiter(bes_ctx *c, double *phi) {
  if (!c || !phi || !c->phi) { set_error("limiter not computed"); return BES_ERR_ARG; }
  return state_to_host(c, c->phi, c->nv, phi);
}

static int fill_bits(bes_ctx *c, unsigned long long *p, int n, unsigned long long v) {
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 64)), dim3(64), 0, c->stream, p, n, v);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

int bes_compute_residual(bes_ctx *c, double *R_out, double *maxabs_out) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  if (!c->R) { int s = dalloc(&c->R, (size_t)c->nv * c->N * c->B); if (s) return s; }
  const int nv = c->nv, B = c->B;
  unsigned long long *bits = c->lbits;
  int s;
  if ((s = reset_err(c)) || (s = fill_bits(c, bits, kLinfSlots * nv * B, 0ull))) return s;
  if ((s = launch_residual(c, nullptr, nullptr, 0.0, 1, c->R, bits))) return s;
  hipLaunchKernelGGL(k_linf_reduce, dim3(grid_for(nv * B, 64)), dim3(64), 0, c->stream, bits, nv * B);
  if (maxabs_out) {
    std::vector<unsigned long long> b(nv * B);
    BES_HIP(hipMemcpyAsync(b.data(), bits, b.size() * 8, hipMemcpyDeviceToHost, c->stream));
    BES_HIP(hipStreamSynchronize(c->stream));
    for (int q = 0; q < B; q++)
      for (int v = 0; v < nv; v++) std::memcpy(&maxabs_out[q * nv + v], &b[v * B + q], 8);
  }
  if (R_out && (s = state_to_host(c, c->R, c->nv, R_out))) return s;
  return check_err(c);
}

// one pseudo-time step (SolverExplicit.cpp:259-393) on the context stream;
// it_host < 0: the Linf row index comes from the device counter
static int explicit_iteration(bes_ctx *c, int nstage, const double *alpha, double cfl, int local, bool linf,
                              int it_host) {
  const int nv = c->nv, B = c->B;
  unsigned long long *bits = c->lbits;
  int s;
  if ((s = launch_timestep(c, c->U0, cfl, local))) return s;
  for (int st = 0; st < nstage; st++) {
    const bool last = st == nstage - 1;
    const double *src = st == 0 ? c->U0 : ((st - 1) % 2 == 0 ? c->Ua : c->Ub);
    double *dst = (last && nstage > 1) ? c->U0 : (st % 2 == 0 ? c->Ua : c->Ub);
    if ((s = launch_gradlim(c, src, nullptr, nullptr))) return s;
    if (last && linf && (s = fill_bits(c, bits, kLinfSlots * nv * B, 0ull))) return s;
    if ((s = launch_residual(c, c->U0, dst, alpha[st], local, nullptr, (last && linf) ? bits : nullptr))) return s;
    if (last && linf) {
      hipLaunchKernelGGL(k_linf_decode, dim3(1), dim3(nv * B), 0, c->stream, bits, c->linf, nv, B,
                         it_host < 0 ? 0 : it_host, c->linf_cap, it_host < 0 ? c->dit : nullptr);
      BES_HIP(hipGetLastError());
    }
    if (last && nstage == 1) std::swap(c->U0, c->Ua);
  }
  return BES_OK;
}

// capture one pseudo-time step as a graph (stage count > 1: the slots are fixed)
static int capture_iteration(bes_ctx *c, int nstage, const double *alpha, double cfl, int local, bool linf) {
  const int key[5] = {nstage, local, linf ? 1 : 0, c->B, c->lim_epoch};
  if (c->graph && std::equal(key, key + 5, c->graph_key) && c->graph_cfl == cfl &&
      std::equal(alpha, alpha + nstage, c->graph_alpha.begin(), c->graph_alpha.end()))
    return BES_OK;
  drop_graph(c);
  const bool prof = c->prof;
  c->prof = false;   // no event records inside the capture
  hipGraph_t g = nullptr;
  BES_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int s = explicit_iteration(c, nstage, alpha, cfl, local, linf, -1);
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  c->prof = prof;
  if (s) { if (g) hipGraphDestroy(g); return s; }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) { c->graph = nullptr; return hip_fail(e, "hipGraphInstantiate"); }
  std::copy(key, key + 5, c->graph_key);
  c->graph_cfl = cfl;
  c->graph_alpha.assign(alpha, alpha + nstage);
  return BES_OK;
}

// every pass of niter pseudo-time steps in one cooperative launch (one
// condition, local time steps; k_explicit_persistent)
static int explicit_run_persistent(bes_ctx *c, int niter, int nstage, const double *alpha, double cfl, bool linf) {
  PersistStep a;
  a.U0 = c->U0;
  a.Ua = c->Ua;
  a.Ub = c->Ub;
  a.PK = c->LG;
  a.FF = c->FF;
  a.dt = c->dt;
  a.frz = c->lim_freeze_on ? c->phi_frz : nullptr;
  a.linf = linf ? c->linf : nullptr;
  a.finf = c->dfinf;
  a.lbits = c->lbits;
  a.err = c->err;
  a.cfl = cfl;
  for (int q = 0; q < 16; q++) a.alpha[q] = q < nstage ? alpha[q] : 0.0;
  a.nstage = nstage;
  a.niter = niter;
  a.linf_cap = c->linf_cap;
  CoefB fm;
  for (int b = 0; b < bes_ctx::kMaxBatch; b++)
    fm.c[b] = (b < (int)c->lim_frozen.size() && c->lim_frozen[b]) ? 1.0 : 0.0;
  const void *fn = c->nd == 2 ? (const void *)k_explicit_persistent<2> : (const void *)k_explicit_persistent<3>;
  const int bs = 256;
  const size_t lds = (size_t)bs * c->nv * (c->nd + 1) * sizeof(double);
  int dev = 0, ncu = 0, per_cu = 0;
  BES_HIP(hipGetDevice(&dev));
  BES_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  BES_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, bs, lds));
  if (per_cu < 1) { set_error("persistent explicit step: no resident block fits"); return BES_ERR_ARG; }
  const int want = std::max(grid_for(c->N, bs), grid_for(c->dm.F, bs));
  const int grid = std::max(1, std::min(per_cu * ncu, want));   // all blocks resident: the barriers complete
  void *args[] = {&a, &c->dm, &c->gc, &fm};
  BES_HIP(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(bs), args, (unsigned)lds, c->stream));
  if (nstage == 1 && niter % 2 == 1) std::swap(c->U0, c->Ua);
  return BES_OK;
}

int bes_explicit_persistent(bes_ctx *c, int on) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  c->use_persistent = on != 0;
  return BES_OK;
}

int bes_explicit_graph(bes_ctx *c, int on) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  c->use_graph = on != 0;
  if (!on) drop_graph(c);
  return BES_OK;
}

int bes_explicit_run(bes_ctx *c, int niter, int nstage, const double *alpha, double cfl, int local,
                     double *linf_out) {
  if (!c || niter < 0 || nstage < 1 || nstage > 16 || !alpha) { set_error("invalid argument"); return BES_ERR_ARG; }
  const int nv = c->nv, B = c->B;
  if (linf_out && c->linf_cap < niter) {
    if (c->linf) hipFree(c->linf);
    c->linf = nullptr;
    c->linf_cap = 0;
    drop_graph(c);   // the capture holds the old Linf buffer
    int s = dalloc(&c->linf, (size_t)niter * nv * B);
    if (s) return s;
    c->linf_cap = niter;
  }
  int s;
  if ((s = reset_err(c))) return s;
  c->dt_local = local != 0;
  const bool persist = c->use_persistent && c->B == 1 && local && !c->prof;
  const bool graph = c->use_graph && nstage > 1 && !c->prof;
  if (persist) {
    if ((s = explicit_run_persistent(c, niter, nstage, alpha, cfl, linf_out != nullptr))) return s;
  } else if (graph) {
    if ((s = capture_iteration(c, nstage, alpha, cfl, local, linf_out != nullptr))) return s;
    BES_HIP(hipMemsetAsync(c->dit, 0, sizeof(int), c->stream));
    for (int it = 0; it < niter; it++) BES_HIP(hipGraphLaunch(c->graph, c->stream));
  } else {
    for (int it = 0; it < niter; it++)
      if ((s = explicit_iteration(c, nstage, alpha, cfl, local, linf_out != nullptr, it))) return s;
  }
  if ((s = prof_collect(c))) return s;
  if (linf_out && niter > 0) {
    BES_HIP(hipMemcpyAsync(linf_out, c->linf, (size_t)niter * nv * B * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  }
  BES_HIP(hipStreamSynchronize(c->stream));
  return check_err(c);
}

// Staging of bes_flux_batch_noalloc: a file-scope device array, one copy per
// GPU for the life of the process (kNoAllocMax faces: UL [n][5], UR [n][6],
// normals [n][3], F [n][5]; 2.5 MB).  Host code never passes the array's name
// to a kernel or a copy: it takes the device address from hipGetSymbolAddress
// once per device, with that device current, and uses that address.
constexpr size_t kNoAllocMax = BES_NOALLOC_MAX_FACES;
static __device__ __align__(256) double scratch[kNoAllocMax * (5 + 6 + 3 + 5)];

__global__ void k_zero_f64(double *p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

static int noalloc_scratch(double **out) {
  constexpr int kMaxDevices = 64;
  static double *ptr[kMaxDevices] = {nullptr};
  int dev = 0;
  BES_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) { set_error("device index out of range"); return BES_ERR_ARG; }
  if (!ptr[dev]) {
    void *p = nullptr;
    BES_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptr[dev] = (double *)p;
  }
  *out = ptr[dev];
  return BES_OK;
}

static int batch_common(int ndim, int kind, long n, const double *UL, const double *UR,
                        const double *nrm, const bes_gas *gas, double *out1, double *out2, bool jac) {
  if ((ndim != 2 && ndim != 3) || n < 0 || (kind < 1 || kind > 3) || (n > 0 && (!UL || !nrm || !out1)) ||
      (kind != kSlipWall && n > 0 && !UR)) {
    set_error("invalid batch argument");
    return BES_ERR_ARG;
  }
  if (n == 0) return BES_OK;
  const int nv = ndim + 2;
  GasConst gc = make_gas(gas);
  std::vector<double> ur((size_t)n * (nv + 1), 0.0);
  if (UR) std::copy(UR, UR + (size_t)n * (nv + 1), ur.begin());
  double *dUL, *dUR, *dn, *d1, *d2 = nullptr;
  const size_t o1 = jac ? (size_t)n * nv * nv : (size_t)n * nv;
  BES_HIP(hipMalloc(&dUL, (size_t)n * nv * 8));
  BES_HIP(hipMalloc(&dUR, (size_t)n * (nv + 1) * 8));
  BES_HIP(hipMalloc(&dn, (size_t)n * ndim * 8));
  BES_HIP(hipMalloc(&d1, o1 * 8));
  if (jac) BES_HIP(hipMalloc(&d2, o1 * 8));
  BES_HIP(hipMemcpy(dUL, UL, (size_t)n * nv * 8, hipMemcpyHostToDevice));
  BES_HIP(hipMemcpy(dUR, ur.data(), (size_t)n * (nv + 1) * 8, hipMemcpyHostToDevice));
  BES_HIP(hipMemcpy(dn, nrm, (size_t)n * ndim * 8, hipMemcpyHostToDevice));
  if (jac) BES_HIP(hipMemset(d2, 0, o1 * 8));
  const int bs = 128;
  if (!jac) {
    const int nb = grid_for((int)n, bs);
    if (ndim == 2) hipLaunchKernelGGL(k_flux_batch<2>, dim3(nb), dim3(bs), 0, 0, kind, (int)n, dUL, dUR, dn, gc, d1);
    else hipLaunchKernelGGL(k_flux_batch<3>, dim3(nb), dim3(bs), 0, 0, kind, (int)n, dUL, dUR, dn, gc, d1);
  } else {
    const int nb = grid_for((int)n * 2 * nv, bs);
    if (ndim == 2) hipLaunchKernelGGL(k_jacobian_batch<2>, dim3(nb), dim3(bs), 0, 0, kind, (int)n, dUL, dUR, dn, gc, d1, d2);
    else hipLaunchKernelGGL(k_jacobian_batch<3>, dim3(nb), dim3(bs), 0, 0, kind, (int)n, dUL, dUR, dn, gc, d1, d2);
  }
  BES_HIP(hipGetLastError());
  BES_HIP(hipMemcpy(out1, d1, o1 * 8, hipMemcpyDeviceToHost));
  if (jac && out2) BES_HIP(hipMemcpy(out2, d2, o1 * 8, hipMemcpyDeviceToHost));
  hipFree(dUL); hipFree(dUR); hipFree(dn); hipFree(d1);
  if (d2) hipFree(d2);
  return BES_OK;
}

int bes_flux_batch(int ndim, int kind, long n, const double *UL, const double *UR, const double *normals,
                   const bes_gas *gas, double *F) {
  return batch_common(ndim, kind, n, UL, UR, normals, gas, F, nullptr, false);
}

int bes_flux_batch_noalloc(int ndim, int kind, long n, const double *UL, const double *UR, const double *normals,
                           const bes_gas *gas, double *F) {
  if ((ndim != 2 && ndim != 3) || n < 0 || (kind < 1 || kind > 3) || (n > 0 && (!UL || !normals || !F)) ||
      (kind != kSlipWall && n > 0 && !UR)) {
    set_error("invalid batch argument");
    return BES_ERR_ARG;
  }
  if (n > BES_NOALLOC_MAX_FACES) {
    set_error("bes_flux_batch_noalloc: n = " + std::to_string(n) + " exceeds the staging limit of " +
              std::to_string(BES_NOALLOC_MAX_FACES) + " faces");
    return BES_ERR_ARG;
  }
  if (n == 0) return BES_OK;
  double *base = nullptr;
  int s = noalloc_scratch(&base);
  if (s) return s;
  const int nv = ndim + 2;
  const size_t nn = (size_t)n;
  // staging regions, each at the offset the largest batch would give it
  double *dUL = base, *dUR = base + kNoAllocMax * 5, *dn = dUR + kNoAllocMax * 6, *dF = dn + kNoAllocMax * 3;
  BES_HIP(hipMemcpy(dUL, UL, nn * nv * 8, hipMemcpyHostToDevice));
  if (UR) {
    BES_HIP(hipMemcpy(dUR, UR, nn * (nv + 1) * 8, hipMemcpyHostToDevice));
  } else {   // slip walls read no right state; zero it so no stale values from an earlier call remain
    const size_t cnt = nn * (nv + 1);
    hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, 0, dUR, cnt);
  }
  BES_HIP(hipMemcpy(dn, normals, nn * ndim * 8, hipMemcpyHostToDevice));
  GasConst gc = make_gas(gas);
  const int bs = 128, nb = grid_for((int)n, bs);
  if (ndim == 2) hipLaunchKernelGGL(k_flux_batch<2>, dim3(nb), dim3(bs), 0, 0, kind, (int)n, dUL, dUR, dn, gc, dF);
  else hipLaunchKernelGGL(k_flux_batch<3>, dim3(nb), dim3(bs), 0, 0, kind, (int)n, dUL, dUR, dn, gc, dF);
  BES_HIP(hipGetLastError());
  BES_HIP(hipMemcpy(F, dF, nn * nv * 8, hipMemcpyDeviceToHost));
  return BES_OK;
}

int bes_jacobian_batch(int ndim, int kind, long n, const double *UL, const double *UR,
                       const double *normals, const bes_gas *gas, double *JL, double *JR) {
  return batch_common(ndim, kind, n, UL, UR, normals, gas, JL, JR, true);
}

long bes_last_error_face(const bes_ctx *c) { return c ? c->last_err_face : -1; }

int bes_set_profiling(bes_ctx *c, int on) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  int s = prof_collect(c);
  if (s) return s;
  c->prof = on != 0;
  for (int q = 0; q < bes_ctx::kProfKinds; q++) { c->prof_ms[q] = 0; c->prof_n[q] = 0; c->prof_bytes[q] = 0; }
  return BES_OK;
}

int bes_get_profile(bes_ctx *c, double *total_ms, long *launches, double *bytes, int nkinds) {
  if (!c || nkinds < 0) { set_error("invalid argument"); return BES_ERR_ARG; }
  int s = prof_collect(c);
  if (s) return s;
  for (int q = 0; q < nkinds && q < bes_ctx::kProfKinds; q++) {
    if (total_ms) total_ms[q] = c->prof_ms[q];
    if (launches) launches[q] = c->prof_n[q];
    if (bytes) bytes[q] = c->prof_bytes[q];
  }
  return BES_OK;
}

int bes_synchronize(bes_ctx *c) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  BES_HIP(hipStreamSynchronize(c->stream));
  return BES_OK;
}

}  // extern "C"
