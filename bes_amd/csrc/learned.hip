// Learned block preconditioner for the implicit path (MI355X, gfx950).
//
// No reference counterpart: the reference solves its implicit systems with
// AMGCL (CoefficientMatrix.cpp:210-325).  This file proposes a cheaper GMRES
// preconditioner with a small graph network and trains it on the GPU on the
// systems the solver assembles.
//
// Operator.  With D_i the diagonal block, O_im the block of the neighbour j
// through local face m and P_im = D_i^-1 O_im D_j^-1,
//     (M v)_i = D_i^-1 v_i - sum_m diag(sigmoid(s_im)) P_im v_j,
// every member lies between block Jacobi (s -> -inf) and the first-order
// Neumann approximation of A^-1 (s -> +inf).  M is stored like A (ELL-BSR,
// Impl::LM) and applied with the same SpMV kernel.
//
// Network (message passing on the cell dual graph, width 32, 3 layers):
//     x_i  = [D_i/dmax_i, b_i/bmax, 1, 0..]           a_im = [O_im/dmax_i, n_im, area/V^(2/3), 0..]
//     h0   = relu(W0 x + c0)
//     msg  = relu(W1 [h_j ; a_im] + c1),  agg_i = sum_m msg_im
//     h   <- h + relu(W2 [h ; agg] + c2)
//     s_im = We [h_i ; h_j] + ce          (nv of 32 outputs)
// Loss: K steps of preconditioned Richardson from probe vectors v (the
// right-hand side and seeded Gaussian vectors), L = mean ||r_K||^2/||v||^2.
//
// Kernels.  Every dense contraction runs on the matrix cores as gfx950's
// V_MFMA_F32_32X32X16_BF16 tiles (bf16 operands, fp32 accumulation): the
// forward layers, the back-propagated products dX = dY W and the weight
// gradients dW = dY^T X (rows of the tile as the reduction dimension).  One
// wave owns 32 cells: a node tile is 32 cells, an edge tile 8 cells x 4 face
// slots.  Lane l of a tile supplies row l%32; lane half h = l/32 supplies
// operand columns [K/2*h, K/2*h + K/2) in 4-element runs (two runs per
// instruction), the same split for both operands, so a lane reads its row (or
// weight row) as one contiguous run.  Output element q of lane l is row
// 8*(q/4) + 4*h + q%4, column l%32.
// Weight gradients are reduced per wave into a partial buffer and summed in a
// fixed order (deterministic).  Adam (Kingma & Ba 2015) updates fp32 master
// weights and refreshes bf16 copies, plain and transposed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bes_hip.h"
#include "context.hpp"
#include "implicit.hpp"
#include "kernels_common.hpp"

namespace bes {

int state_to_device(bes_ctx *c, const double *host, int ncomp, double *dsoa);

namespace lp {

constexpr int WD = 32, NL = 3, SL = 4;       // width, layers, face slots per cell
constexpr int OFF_W0 = 0, OFF_C0 = 1024;
__host__ __device__ constexpr int off_layer(int l) { return 1056 + l * 4160; }
constexpr int REL_W1 = 0, REL_C1 = 2048, REL_W2 = 2080, REL_C2 = 4128;
constexpr int OFF_WE = 13536, OFF_CE = 15584, NPARAM = 15616;
constexpr int kTrainBlocks = 256;            // fixed grid of the backward kernels
constexpr int kTrainWaves = kTrainBlocks * 4;
constexpr int kDotBlocks = 512;

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef short bf8s __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

// float -> bfloat16, round to nearest even
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

union Pack8 {
  uint4 u;
  bf4 v[2];
  uint16_t s[8];
};

// n/4 bf4 operands from n contiguous bf16 (n = 16 or 32), or zeros
template <int NE>
__device__ __forceinline__ void load_run(const uint16_t *p, bool ok, bf4 *o) {
#pragma unroll
  for (int q = 0; q < NE / 8; q++) {
    Pack8 t;
    t.u = ok ? reinterpret_cast<const uint4 *>(p)[q] : make_uint4(0, 0, 0, 0);
    o[2 * q] = t.v[0];
    o[2 * q + 1] = t.v[1];
  }
}

__device__ __forceinline__ f16v zero16() {
  f16v z;
#pragma unroll
  for (int q = 0; q < 16; q++) z[q] = 0.f;
  return z;
}

// KS 4-element operand runs per lane (run s of A and run s of B hold the
// same k indices for every lane) through the 32x32x16 BF16 MFMA, two runs per
// instruction: a lane's 8-element operand is its runs s and s+1 back to back,
// in A and in B alike, so every product pairs the same k of A and B and each
// instruction sums the k of both runs (in its own order)
template <int KS>
__device__ __forceinline__ f16v mma_run(const bf4 *a, const bf4 *b, f16v acc) {
  static_assert(KS % 2 == 0, "operand runs are consumed in pairs");
#pragma unroll
  for (int s = 0; s < KS; s += 2) {
    const bf8s sa = __builtin_shufflevector(a[s], a[s + 1], 0, 1, 2, 3, 4, 5, 6, 7);
    const bf8s sb = __builtin_shufflevector(b[s], b[s + 1], 0, 1, 2, 3, 4, 5, 6, 7);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8, sa), __builtin_bit_cast(bf8, sb), acc, 0,
                                                  0, 0);
  }
  return acc;
}

// column-wise operand of a [32][pitch] bf16 LDS tile: rows 16h + 4s + j, column col
__device__ __forceinline__ void lds_col(const uint16_t *tile, int pitch, int h, int col, bf4 *o) {
#pragma unroll
  for (int s = 0; s < 4; s++) {
    bf4 v;
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = (short)tile[(16 * h + 4 * s + j) * pitch + col];
    o[s] = v;
  }
}

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
template <int ND>
__global__ __launch_bounds__(256) void k_bmax(const double *__restrict__ b, int N,
                                             unsigned long long *__restrict__ bits) {
  constexpr int NV = ND + 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
  for (int v = 0; v < NV; v++) {
    double x = i < N ? ::fabs(b[(size_t)i * NV + v]) : 0.0;
    x = block_max(x);
    if (threadIdx.x == 0) atomicMax(&bits[v], (unsigned long long)__double_as_longlong(x));
  }
}

// node features x, edge features a, and M's diagonal blocks D_i^-1
template <int ND, typename TM>
__global__ __launch_bounds__(256) void k_features(const TM *__restrict__ A, const double *__restrict__ b,
                                                 const unsigned long long *__restrict__ bbits, DevMesh dm,
                                                 const float *__restrict__ efeat, uint16_t *__restrict__ X,
                                                 uint16_t *__restrict__ Aed, TM *__restrict__ LM) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int N = dm.N;
  const int i = cell_index();
  if (i >= N) return;
  double D[NB], Inv[NB];
  load_block<TM, NB>(A, 0, N, i, D);
  double dmax = 0;
#pragma unroll
  for (int e = 0; e < NB; e++) dmax = fmax(dmax, ::fabs(D[e]));
  Pack8 row[4];
#pragma unroll
  for (int q = 0; q < 32; q++) {
    double v = 0.0;
    if (q < NB) v = D[q] / dmax;
    else if (q < NB + NV) {
      double bm = __longlong_as_double((long long)bbits[q - NB]);
      v = b[(size_t)i * NV + (q - NB)] / (bm > 0 ? bm : 1.0);
    } else if (q == NB + NV) v = 1.0;
    row[q >> 3].s[q & 7] = f2bf((float)v);
  }
#pragma unroll
  for (int q = 0; q < 4; q++) reinterpret_cast<uint4 *>(X + (size_t)i * WD)[q] = row[q].u;
#pragma unroll
  for (int m = 0; m < SL; m++) {
    const bool edge = m < K && dm.nbr[(size_t)m * N + i] >= 0;
    double O[NB];
    if (edge) load_block<TM, NB>(A, 1 + m, N, i, O);
#pragma unroll
    for (int q = 0; q < 32; q++) {
      float v = 0.f;
      if (edge) {
        if (q < NB) v = (float)(O[q] / dmax);
        else if (q < NB + ND + 1) v = efeat[((size_t)i * SL + m) * 4 + (q - NB)];
      }
      row[q >> 3].s[q & 7] = f2bf(v);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) reinterpret_cast<uint4 *>(Aed + ((size_t)i * SL + m) * WD)[q] = row[q].u;
  }
  block_inverse<NV>(D, Inv);
  store_block<TM, NB>(LM, 0, N, i, Inv);
}

// P_im = D_i^-1 O_im D_j^-1 (float), zero off the graph
template <int ND, typename TM>
__global__ __launch_bounds__(256) void k_neumann(const TM *__restrict__ A, const TM *__restrict__ LM, DevMesh dm,
                                                 float *__restrict__ P) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int N = dm.N;
  const int i = cell_index();
  if (i >= N) return;
  double Di[NB];
  load_block<TM, NB>(LM, 0, N, i, Di);
#pragma unroll
  for (int m = 0; m < K; m++) {
    const int j = dm.nbr[(size_t)m * N + i];
    double out[NB];
    if (j >= 0) {
      double O[NB], Dj[NB], T[NB];
      load_block<TM, NB>(A, 1 + m, N, i, O);
      load_block<TM, NB>(LM, 0, N, j, Dj);
#pragma unroll
      for (int r = 0; r < NV; r++)
#pragma unroll
        for (int cc = 0; cc < NV; cc++) {
          double s = 0;
#pragma unroll
          for (int q = 0; q < NV; q++) s += Di[r * NV + q] * O[q * NV + cc];
          T[r * NV + cc] = s;
        }
#pragma unroll
      for (int r = 0; r < NV; r++)
#pragma unroll
        for (int cc = 0; cc < NV; cc++) {
          double s = 0;
#pragma unroll
          for (int q = 0; q < NV; q++) s += T[r * NV + q] * Dj[q * NV + cc];
          out[r * NV + cc] = s;
        }
    } else {
#pragma unroll
      for (int e = 0; e < NB; e++) out[e] = 0.0;
    }
#pragma unroll
    for (int e = 0; e < NB; e++) P[((size_t)m * NB + e) * N + i] = (float)out[e];
  }
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
struct Net {
  const uint16_t *wb;   // bf16 weights, flat layout
  const uint16_t *wt;   // bf16 weights transposed per matrix ([in][32])
  const float *theta;   // fp32 master weights (biases read here)
};

// h0 = relu(W0 x + c0); one wave per node tile
__global__ __launch_bounds__(256) void k_encode(const uint16_t *__restrict__ X, Net net, int N,
                                                uint16_t *__restrict__ H0) {
  const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
  const int base = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
  const int node = base + col;
  bf4 a[4], b[4];
  load_run<16>(X + (size_t)node * WD + 16 * h, node < N, a);
  load_run<16>(net.wb + OFF_W0 + col * 32 + 16 * h, true, b);
  f16v acc = mma_run<4>(a, b, zero16());
  const float bias = net.theta[OFF_C0 + col];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int nd = base + 8 * (q >> 2) + 4 * h + (q & 3);
    if (nd < N) H0[(size_t)nd * WD + col] = f2bf(fmaxf(acc[q] + bias, 0.f));
  }
}

// one message-passing layer; one wave per node tile (4 edge tiles + 1 node tile)
__global__ __launch_bounds__(256) void k_layer(int l, const uint16_t *__restrict__ Hin,
                                               const uint16_t *__restrict__ Aed, const int *__restrict__ nbr, int K,
                                               Net net, int N, uint16_t *__restrict__ Msg,
                                               uint16_t *__restrict__ Agg, uint16_t *__restrict__ Rr,
                                               uint16_t *__restrict__ Hout) {
  __shared__ __align__(16) uint16_t aggL[4][32][WD];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
  const int base = (blockIdx.x * 4 + w) * 32;
  const int off = off_layer(l);
  bf4 b1[8], b2[8];
  load_run<32>(net.wb + off + REL_W1 + col * 64 + 32 * h, true, b1);
  load_run<32>(net.wb + off + REL_W2 + col * 64 + 32 * h, true, b2);
  const float c1 = net.theta[off + REL_C1 + col], c2 = net.theta[off + REL_C2 + col];
#pragma unroll 1
  for (int t = 0; t < 4; t++) {
    const int node = base + 8 * t + (col >> 2), slot = col & 3;
    const int j = (node < N && slot < K) ? nbr[(size_t)slot * N + node] : -1;
    bf4 a[8];
    if (h == 0) load_run<32>(Hin + (size_t)(j < 0 ? 0 : j) * WD, j >= 0, a);
    else load_run<32>(Aed + ((size_t)node * SL + slot) * WD, node < N, a);
    f16v acc = mma_run<8>(a, b1, zero16());
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const int nd = base + 8 * t + 2 * g + h;
      float sum = 0.f;
#pragma unroll
      for (int s = 0; s < 4; s++) {
        const bool e = nd < N && s < K && nbr[(size_t)s * N + nd] >= 0;
        const float v = e ? fmaxf(acc[4 * g + s] + c1, 0.f) : 0.f;
        if (nd < N) Msg[((size_t)nd * SL + s) * WD + col] = f2bf(v);
        sum += v;
      }
      aggL[w][8 * t + 2 * g + h][col] = f2bf(sum);
    }
  }
  __syncthreads();
  const int node = base + col;
  bf4 a[8];
  if (h == 0) load_run<32>(Hin + (size_t)node * WD, node < N, a);
  else load_run<32>(&aggL[w][col][0], true, a);
  f16v acc = mma_run<8>(a, b2, zero16());
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int rho = 8 * (q >> 2) + 4 * h + (q & 3), nd = base + rho;
    if (nd < N) {
      const float v = fmaxf(acc[q] + c2, 0.f);
      const size_t o = (size_t)nd * WD + col;
      Rr[o] = f2bf(v);
      Hout[o] = f2bf(bf2f(Hin[o]) + v);
      Agg[o] = aggL[w][rho][col];
    }
  }
}

// blend logits -> G = sigmoid, M_im = -diag(G) P_im
template <int ND, typename TM>
__global__ __launch_bounds__(256) void k_decode(const uint16_t *__restrict__ H, DevMesh dm, Net net,
                                               const float *__restrict__ P, float *__restrict__ Gb,
                                               TM *__restrict__ LM) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int N = dm.N;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
  const int base = (blockIdx.x * 4 + w) * 32;
  bf4 b[8];
  load_run<32>(net.wb + OFF_WE + col * 64 + 32 * h, true, b);
  const float ce = net.theta[OFF_CE + col];
#pragma unroll 1
  for (int t = 0; t < 4; t++) {
    const int node = base + 8 * t + (col >> 2), slot = col & 3;
    const int j = (node < N && slot < K) ? dm.nbr[(size_t)slot * N + node] : -1;
    bf4 a[8];
    if (h == 0) load_run<32>(H + (size_t)node * WD, node < N, a);
    else load_run<32>(H + (size_t)(j < 0 ? 0 : j) * WD, j >= 0, a);
    f16v acc = mma_run<8>(a, b, zero16());
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int nd = base + 8 * t + 2 * (q >> 2) + h, s = q & 3;
      if (nd >= N || s >= K || col >= NV) continue;
      const bool e = dm.nbr[(size_t)s * N + nd] >= 0;
      const float G = e ? 1.f / (1.f + __expf(-(acc[q] + ce))) : 0.f;
      Gb[((size_t)nd * SL + s) * 8 + col] = G;
#pragma unroll
      for (int cc = 0; cc < NV; cc++) {
        const size_t e_ = (size_t)(s * NB + col * NV + cc);
        LM[((size_t)(1 + s) * NB + col * NV + cc) * N + nd] = (TM)(e ? -G * P[e_ * N + nd] : 0.f);
      }
    }
  }
}

// FP16 copy of the operator for the FP32 solves
__global__ void k_to_half(const float *__restrict__ x, __half *__restrict__ y, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    y[q] = __float2half(x[q]);
}

// ---------------------------------------------------------------------------
// loss: Richardson chains and their adjoint (vectors [N][nv], float)
// ---------------------------------------------------------------------------
// (A^T x)_j = A_jj^T x_j + sum_m A_(i, 1+mrev(j,m))^T x_i, i = nbr(j,m)
template <int ND, typename TM, typename TV>
__global__ __launch_bounds__(256) void k_spmv_t(const TM *__restrict__ A, DevMesh dm, const int *__restrict__ mrev,
                                                const TV *__restrict__ x, TV *__restrict__ y) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int N = dm.N;
  const int j = cell_index();
  if (j >= N) return;
  double acc[NV], xv[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) xv[c] = (double)x[(size_t)j * NV + c];
#pragma unroll
  for (int r = 0; r < NV; r++) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < NV; c++) s += (double)A[((size_t)(c * NV + r)) * N + j] * xv[c];
    acc[r] = s;
  }
#pragma unroll
  for (int m = 0; m < K; m++) {
    const int i = dm.nbr[(size_t)m * N + j];
    if (i < 0) continue;
    const int mm = mrev[(size_t)m * N + j];
#pragma unroll
    for (int c = 0; c < NV; c++) xv[c] = (double)x[(size_t)i * NV + c];
    const TM *Ao = A + (size_t)(1 + mm) * NB * N;
#pragma unroll
    for (int r = 0; r < NV; r++) {
      double s = 0;
#pragma unroll
      for (int c = 0; c < NV; c++) s += (double)Ao[((size_t)(c * NV + r)) * N + i] * xv[c];
      acc[r] += s;
    }
  }
#pragma unroll
  for (int r = 0; r < NV; r++) y[(size_t)j * NV + r] = (TV)acc[r];
}

// out = x - y (r_{k+1} = r_k - A M r_k), or out = x - y scaled
__global__ void k_sub(const float *__restrict__ x, const float *__restrict__ y, float *__restrict__ out, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    out[q] = x[q] - y[q];
}
__global__ void k_scale_f(const float *__restrict__ x, float a, float *__restrict__ out, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    out[q] = a * x[q];
}
__global__ void k_from_double(const double *__restrict__ x, float *__restrict__ out, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    out[q] = (float)x[q];
}
__global__ void k_fill_f(float *__restrict__ p, float v, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    p[q] = v;
}

// seeded standard normal vector (splitmix64 counter stream, Box-Muller)
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__global__ void k_randn(unsigned long long seed, float *__restrict__ out, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const unsigned long long z = mix64(seed ^ mix64(q));
    const double u1 = ((z >> 11) + 1.0) * (1.0 / 9007199254740993.0);
    const double u2 = (mix64(z) >> 11) * (1.0 / 9007199254740992.0);
    out[q] = (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
  }
}

// sum of squares, per-block partials (double)
__global__ void k_sumsq(const float *__restrict__ x, size_t n, double *__restrict__ part) {
  double s = 0;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    s += (double)x[q] * x[q];
  __shared__ double sh[4];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
__global__ void k_sum_parts(const double *__restrict__ part, int nblk, double *__restrict__ out) {
  double s = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[b];
  __shared__ double sh[4];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = sh[0] + sh[1] + sh[2] + sh[3];
}

// dMe_im -= at_i (x) r_j   (at = A^T g, one (probe, step) term)
template <int ND>
__global__ __launch_bounds__(256) void k_dme(const float *__restrict__ at, const float *__restrict__ rk, DevMesh dm,
                                             float *__restrict__ dMe) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int N = dm.N;
  const int i = cell_index();
  if (i >= N) return;
  float ai[NV];
#pragma unroll
  for (int r = 0; r < NV; r++) ai[r] = at[(size_t)i * NV + r];
#pragma unroll
  for (int m = 0; m < K; m++) {
    const int j = dm.nbr[(size_t)m * N + i];
    if (j < 0) continue;
    float rj[NV];
#pragma unroll
    for (int cc = 0; cc < NV; cc++) rj[cc] = rk[(size_t)j * NV + cc];
#pragma unroll
    for (int r = 0; r < NV; r++)
#pragma unroll
      for (int cc = 0; cc < NV; cc++) dMe[((size_t)m * NB + r * NV + cc) * N + i] -= ai[r] * rj[cc];
  }
}

// dL/ds_im = -sum_c dMe_im[r][c] P_im[r][c] * G (1 - G), as bf16 rows of 32
template <int ND>
__global__ __launch_bounds__(256) void k_dlogit(const float *__restrict__ dMe, const float *__restrict__ P,
                                                const float *__restrict__ Gb, DevMesh dm,
                                                uint16_t *__restrict__ dlog) {
  constexpr int NV = ND + 2, K = ND + 1, NB = NV * NV;
  const int N = dm.N;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t >> 2, m = t & 3;
  if (i >= N) return;
  const bool e = m < K && dm.nbr[(size_t)m * N + i] >= 0;
  Pack8 row[4];
#pragma unroll
  for (int q = 0; q < 32; q++) row[q >> 3].s[q & 7] = 0;
  if (e) {
#pragma unroll
    for (int r = 0; r < NV; r++) {
      float s = 0.f;
#pragma unroll
      for (int cc = 0; cc < NV; cc++) {
        const size_t q = ((size_t)m * NB + r * NV + cc) * N + i;
        s += dMe[q] * P[q];
      }
      const float G = Gb[((size_t)i * SL + m) * 8 + r];
      row[r >> 3].s[r & 7] = f2bf(-s * G * (1.f - G));
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) reinterpret_cast<uint4 *>(dlog + ((size_t)i * SL + m) * WD)[q] = row[q].u;
}

// ---------------------------------------------------------------------------
// backward (fixed grid kTrainBlocks x 4 waves; every wave writes its whole
// partial-gradient range, zeros when it had no tile)
// ---------------------------------------------------------------------------
// store a 32 x (32*NT) weight-gradient tile set and its bias sum into this wave's partial row
template <int NT>
__device__ __forceinline__ void store_wgrad(float *__restrict__ row, int woff, int boff, const f16v *acc,
                                            float bsum, int h, int col) {
#pragma unroll
  for (int tt = 0; tt < NT; tt++)
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int o = 8 * (q >> 2) + 4 * h + (q & 3);
      row[woff + o * (32 * NT) + 32 * tt + col] = acc[tt][q];
    }
  bsum += __shfl_xor(bsum, 32, 64);
  if (h == 0) row[boff + col] = bsum;
}

__global__ __launch_bounds__(256) void k_decode_bwd(const uint16_t *__restrict__ dlog, const uint16_t *__restrict__ H,
                                                    const int *__restrict__ nbr, int K, Net net, int N, int ntile,
                                                    float *__restrict__ dA, float *__restrict__ dnb,
                                                    float *__restrict__ gpart) {
  __shared__ __align__(16) uint16_t dL[4][32][WD];
  __shared__ __align__(16) uint16_t pL[4][32][2 * WD];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
  bf4 b0[4], b1[4];
  load_run<16>(net.wt + OFF_WE + col * 32 + 16 * h, true, b0);
  load_run<16>(net.wt + OFF_WE + (32 + col) * 32 + 16 * h, true, b1);
  f16v g[2] = {zero16(), zero16()};
  float gb = 0.f;
  for (int bt = blockIdx.x; bt * 4 < ntile; bt += gridDim.x) {
    const int base = (bt * 4 + w) * 32;
#pragma unroll 1
    for (int t = 0; t < 4; t++) {
      const int node = base + 8 * t + (col >> 2), slot = col & 3;
      const int j = (node < N && slot < K) ? nbr[(size_t)slot * N + node] : -1;
      bf4 a[4];
      load_run<16>(dlog + ((size_t)node * SL + slot) * WD + 16 * h, node < N, a);
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int q = 0; q < 4; q++) dL[w][col][16 * h + 4 * s + q] = (uint16_t)a[s][q];
      {
        bf4 p[8];
        if (h == 0) load_run<32>(H + (size_t)node * WD, node < N, p);
        else load_run<32>(H + (size_t)(j < 0 ? 0 : j) * WD, j >= 0, p);
#pragma unroll
        for (int s = 0; s < 8; s++)
#pragma unroll
          for (int q = 0; q < 4; q++) pL[w][col][32 * h + 4 * s + q] = (uint16_t)p[s][q];
      }
      f16v acc0 = mma_run<4>(a, b0, zero16());
      f16v acc1 = mma_run<4>(a, b1, zero16());
#pragma unroll
      for (int gg = 0; gg < 4; gg++) {
        const int nd = base + 8 * t + 2 * gg + h;
        if (nd >= N) continue;
        dA[(size_t)nd * WD + col] = acc0[4 * gg] + acc0[4 * gg + 1] + acc0[4 * gg + 2] + acc0[4 * gg + 3];
#pragma unroll
        for (int s = 0; s < 4; s++) dnb[((size_t)nd * SL + s) * WD + col] = acc1[4 * gg + s];
      }
      __syncthreads();
      bf4 ad[4], bp0[4], bp1[4];
      lds_col(&dL[w][0][0], WD, h, col, ad);
      lds_col(&pL[w][0][0], 2 * WD, h, col, bp0);
      lds_col(&pL[w][0][0], 2 * WD, h, 32 + col, bp1);
      g[0] = mma_run<4>(ad, bp0, g[0]);
      g[1] = mma_run<4>(ad, bp1, g[1]);
#pragma unroll
      for (int r = 0; r < 16; r++) gb += bf2f(dL[w][16 * h + r][col]);
      __syncthreads();
    }
  }
  store_wgrad<2>(gpart + (size_t)(blockIdx.x * 4 + w) * NPARAM, OFF_WE, OFF_CE, g, gb, h, col);
}

// out[i] = in[i] + sum_m back[nbr(i,m)][mrev(i,m)]
__global__ __launch_bounds__(256) void k_gather_rev(const float *__restrict__ in, const float *__restrict__ back,
                                                    const int *__restrict__ nbr, const int *__restrict__ mrev, int K,
                                                    int N, float *__restrict__ out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = (int)(t >> 5), col = (int)(t & 31);
  if (i >= N) return;
  float v = in[(size_t)i * WD + col];
  for (int m = 0; m < K; m++) {
    const int j = nbr[(size_t)m * N + i];
    if (j >= 0) v += back[((size_t)j * SL + mrev[(size_t)m * N + i]) * WD + col];
  }
  out[(size_t)i * WD + col] = v;
}

// backward of one layer: node part (W2) then 4 edge tiles (W1)
__global__ __launch_bounds__(256) void k_layer_bwd(int l, const float *__restrict__ dH, const uint16_t *__restrict__ Hl,
                                                   const uint16_t *__restrict__ Agg, const uint16_t *__restrict__ Rr,
                                                   const uint16_t *__restrict__ Msg, const uint16_t *__restrict__ Aed,
                                                   const int *__restrict__ nbr, int K, Net net, int N, int ntile,
                                                   float *__restrict__ dHacc, float *__restrict__ dnb,
                                                   float *__restrict__ gpart) {
  __shared__ __align__(16) uint16_t dL[4][32][WD];
  __shared__ __align__(16) uint16_t xL[4][32][2 * WD];
  __shared__ __align__(16) float gaL[4][32][WD];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
  const int off = off_layer(l);
  bf4 w2a[4], w2b[4], w1a[4];
  load_run<16>(net.wt + off + REL_W2 + col * 32 + 16 * h, true, w2a);
  load_run<16>(net.wt + off + REL_W2 + (32 + col) * 32 + 16 * h, true, w2b);
  load_run<16>(net.wt + off + REL_W1 + col * 32 + 16 * h, true, w1a);
  f16v g2[2] = {zero16(), zero16()}, g1[2] = {zero16(), zero16()};
  float gb2 = 0.f, gb1 = 0.f;
  for (int bt = blockIdx.x; bt * 4 < ntile; bt += gridDim.x) {
    const int base = (bt * 4 + w) * 32;
    {  // node tile: dpre2 = bf16(dH * (r > 0)), du = dpre2 W2
      const int node = base + col;
      const bool ok = node < N;
      bf4 a[4];
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int kk = 16 * h + 4 * s + q;
          float d = 0.f;
          if (ok && bf2f(Rr[(size_t)node * WD + kk]) > 0.f) d = dH[(size_t)node * WD + kk];
          const uint16_t bv = f2bf(d);
          a[s][q] = (short)bv;
          dL[w][col][kk] = bv;
        }
      {
        bf4 u[8];
        if (h == 0) load_run<32>(Hl + (size_t)node * WD, ok, u);
        else load_run<32>(Agg + (size_t)node * WD, ok, u);
#pragma unroll
        for (int s = 0; s < 8; s++)
#pragma unroll
          for (int q = 0; q < 4; q++) xL[w][col][32 * h + 4 * s + q] = (uint16_t)u[s][q];
      }
      f16v acc0 = mma_run<4>(a, w2a, zero16());
      f16v acc1 = mma_run<4>(a, w2b, zero16());
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int rho = 8 * (q >> 2) + 4 * h + (q & 3), nd = base + rho;
        if (nd < N) dHacc[(size_t)nd * WD + col] = dH[(size_t)nd * WD + col] + acc0[q];
        gaL[w][rho][col] = acc1[q];
      }
      __syncthreads();
      bf4 ad[4], bx0[4], bx1[4];
      lds_col(&dL[w][0][0], WD, h, col, ad);
      lds_col(&xL[w][0][0], 2 * WD, h, col, bx0);
      lds_col(&xL[w][0][0], 2 * WD, h, 32 + col, bx1);
      g2[0] = mma_run<4>(ad, bx0, g2[0]);
      g2[1] = mma_run<4>(ad, bx1, g2[1]);
#pragma unroll
      for (int r = 0; r < 16; r++) gb2 += bf2f(dL[w][16 * h + r][col]);
      __syncthreads();
    }
#pragma unroll 1
    for (int t = 0; t < 4; t++) {  // edge tiles: dpre1 = bf16(dagg * (msg > 0)), dinp = dpre1 W1
      const int ln = 8 * t + (col >> 2), node = base + ln, slot = col & 3;
      const bool ok = node < N;
      const int j = (ok && slot < K) ? nbr[(size_t)slot * N + node] : -1;
      bf4 a[4];
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int kk = 16 * h + 4 * s + q;
          float d = 0.f;
          if (ok && bf2f(Msg[((size_t)node * SL + slot) * WD + kk]) > 0.f) d = gaL[w][ln][kk];
          const uint16_t bv = f2bf(d);
          a[s][q] = (short)bv;
          dL[w][col][kk] = bv;
        }
      {
        bf4 u[8];
        if (h == 0) load_run<32>(Hl + (size_t)(j < 0 ? 0 : j) * WD, j >= 0, u);
        else load_run<32>(Aed + ((size_t)node * SL + slot) * WD, ok, u);
#pragma unroll
        for (int s = 0; s < 8; s++)
#pragma unroll
          for (int q = 0; q < 4; q++) xL[w][col][32 * h + 4 * s + q] = (uint16_t)u[s][q];
      }
      f16v acc = mma_run<4>(a, w1a, zero16());
#pragma unroll
      for (int gg = 0; gg < 4; gg++) {
        const int nd = base + 8 * t + 2 * gg + h;
        if (nd >= N) continue;
#pragma unroll
        for (int s = 0; s < 4; s++) dnb[((size_t)nd * SL + s) * WD + col] = acc[4 * gg + s];
      }
      __syncthreads();
      bf4 ad[4], bx0[4], bx1[4];
      lds_col(&dL[w][0][0], WD, h, col, ad);
      lds_col(&xL[w][0][0], 2 * WD, h, col, bx0);
      lds_col(&xL[w][0][0], 2 * WD, h, 32 + col, bx1);
      g1[0] = mma_run<4>(ad, bx0, g1[0]);
      g1[1] = mma_run<4>(ad, bx1, g1[1]);
#pragma unroll
      for (int r = 0; r < 16; r++) gb1 += bf2f(dL[w][16 * h + r][col]);
      __syncthreads();
    }
  }
  float *row = gpart + (size_t)(blockIdx.x * 4 + w) * NPARAM;
  store_wgrad<2>(row, off + REL_W2, off + REL_C2, g2, gb2, h, col);
  store_wgrad<2>(row, off + REL_W1, off + REL_C1, g1, gb1, h, col);
}

__global__ __launch_bounds__(256) void k_encode_bwd(const float *__restrict__ dH, const uint16_t *__restrict__ H0,
                                                    const uint16_t *__restrict__ X, int N, int ntile,
                                                    float *__restrict__ gpart) {
  __shared__ __align__(16) uint16_t dL[4][32][WD];
  __shared__ __align__(16) uint16_t xL[4][32][WD];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
  f16v g[1] = {zero16()};
  float gb = 0.f;
  for (int bt = blockIdx.x; bt * 4 < ntile; bt += gridDim.x) {
    const int base = (bt * 4 + w) * 32, node = base + col;
    const bool ok = node < N;
#pragma unroll
    for (int kk0 = 0; kk0 < 16; kk0++) {
      const int kk = 16 * h + kk0;
      float d = 0.f;
      if (ok && bf2f(H0[(size_t)node * WD + kk]) > 0.f) d = dH[(size_t)node * WD + kk];
      dL[w][col][kk] = f2bf(d);
      xL[w][col][kk] = ok ? X[(size_t)node * WD + kk] : (uint16_t)0;
    }
    __syncthreads();
    bf4 ad[4], bx[4];
    lds_col(&dL[w][0][0], WD, h, col, ad);
    lds_col(&xL[w][0][0], WD, h, col, bx);
    g[0] = mma_run<4>(ad, bx, g[0]);
#pragma unroll
    for (int r = 0; r < 16; r++) gb += bf2f(dL[w][16 * h + r][col]);
    __syncthreads();
  }
  store_wgrad<1>(gpart + (size_t)(blockIdx.x * 4 + w) * NPARAM, OFF_W0, OFF_C0, g, gb, h, col);
}

__global__ void k_grad_reduce(const float *__restrict__ gpart, int nrow, float *__restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float s = 0.f;
  for (int r = 0; r < nrow; r++) s += gpart[(size_t)r * NPARAM + p];
  grad[p] = s;
}

// which weight matrix holds parameter p (-1: a bias); in-count of that matrix
__device__ __forceinline__ int matrix_of(int p, int *in_count) {
  if (p < OFF_C0) { *in_count = 32; return OFF_W0; }
  if (p < off_layer(0)) return -1;
  if (p < OFF_WE) {
    const int l = (p - off_layer(0)) / 4160, rel = (p - off_layer(0)) % 4160, o = off_layer(l);
    *in_count = 64;
    if (rel < REL_C1) return o + REL_W1;
    if (rel >= REL_W2 && rel < REL_C2) return o + REL_W2;
    return -1;
  }
  if (p < OFF_CE) { *in_count = 64; return OFF_WE; }
  return -1;
}

// Adam step (lr > 0) and refresh of the bf16 copies
__global__ void k_adam(float *__restrict__ theta, const float *__restrict__ grad, float *__restrict__ m1,
                       float *__restrict__ m2, float lr, float c1, float c2, uint16_t *__restrict__ wb,
                       uint16_t *__restrict__ wt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float th = theta[p];
  if (lr > 0.f) {
    const float g = grad[p];
    const float a = 0.9f * m1[p] + 0.1f * g, b = 0.999f * m2[p] + 0.001f * g * g;
    m1[p] = a;
    m2[p] = b;
    th -= lr * (a / c1) / (sqrtf(b / c2) + 1e-8f);
    theta[p] = th;
  }
  const uint16_t v = f2bf(th);
  wb[p] = v;
  int in;
  const int off = matrix_of(p, &in);
  if (off >= 0) {
    const int o = (p - off) / in, i = (p - off) % in;
    wt[off + i * 32 + o] = v;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct Lp {
  int N = 0, k = 0, nv = 0, nd = 0, ntile = 0;
  float *theta = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr, *gpart = nullptr;
  uint16_t *wb = nullptr, *wt = nullptr;
  int adam_t = 0;
  unsigned long long *bbits = nullptr;   // [8]
  float *efeat = nullptr;                // [N][4][4]
  uint16_t *X = nullptr, *Aed = nullptr, *H = nullptr, *Msg = nullptr, *Agg = nullptr, *Rr = nullptr;
  float *P = nullptr, *G = nullptr;
  // training
  float *dH = nullptr, *dH2 = nullptr, *dA = nullptr, *dnb = nullptr, *dMe = nullptr;
  uint16_t *dlog = nullptr;
  float *chain = nullptr;                // [Q][K+1][n]
  float *tv1 = nullptr, *tv2 = nullptr, *gv = nullptr;
  double *part = nullptr, *dsum = nullptr;
  int qcap = 0, kcap = 0;
  bool train_ready = false;
};

template <typename T>
static int lalloc(T **p, size_t n) {
  BES_HIP(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
  return BES_OK;
}

static Lp *lp_of(Impl *im) { return (Lp *)im->lp; }

static uint64_t splitmix(uint64_t &state) {
  state += 0x9E3779B97F4A7C15ull;
  uint64_t z = state;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// deterministic initial parameters (same stream as the CPU restatement)
static std::vector<float> init_params(uint64_t seed, int nv) {
  std::vector<float> p(NPARAM, 0.f);
  uint64_t st = seed;
  auto fill = [&](int off, int out, int in, double scale) {
    const double lim = std::sqrt(6.0 / (in + out)) * scale;
    for (int q = 0; q < out * in; q++) {
      const uint64_t z = splitmix(st);
      p[off + q] = (float)(((double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * lim);
    }
  };
  fill(OFF_W0, 32, 32, 1.0);
  for (int l = 0; l < NL; l++) {
    fill(off_layer(l) + REL_W1, 32, 64, 1.0);
    fill(off_layer(l) + REL_W2, 32, 64, 0.5);
  }
  const int used = nv * nv + nv + 1;
  for (int o = 0; o < 32; o++)
    for (int i = used; i < 32; i++) p[OFF_W0 + o * 32 + i] = 0.f;
  for (int r = 0; r < nv; r++) p[OFF_CE + r] = 3.0f;   // initial blend sigmoid(3) = 0.953
  return p;
}

static int lp_refresh(bes_ctx *c, Lp *L) {
  hipLaunchKernelGGL(k_adam, dim3(grid_for(NPARAM, 256)), dim3(256), 0, c->stream, L->theta, L->grad, L->m1, L->m2,
                     0.f, 1.f, 1.f, L->wb, L->wt);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

int lp_release(Impl *im) {
  Lp *L = lp_of(im);
  if (!L) return BES_OK;
  void *ptrs[] = {L->theta, L->grad, L->m1, L->m2, L->gpart, L->wb, L->wt, L->bbits, L->efeat, L->X, L->Aed, L->H,
                  L->Msg, L->Agg, L->Rr, L->P, L->G, L->dH, L->dH2, L->dA, L->dnb, L->dMe, L->dlog, L->chain,
                  L->tv1, L->tv2, L->gv, L->part, L->dsum};
  for (void *p : ptrs)
    if (p) hipFree(p);
  delete L;
  im->lp = nullptr;
  return BES_OK;
}

static int lp_alloc(bes_ctx *c, Impl *im, uint64_t seed) {
  lp_release(im);
  Lp *L = new Lp();
  im->lp = L;
  const int N = c->N, k = c->k, nv = c->nv, nd = c->nd, NB = nv * nv;
  L->N = N; L->k = k; L->nv = nv; L->nd = nd;
  L->ntile = (N + 31) / 32;
  const size_t Np = (size_t)L->ntile * 32;   // padded row count of the tiles
  int s;
  if ((s = lalloc(&L->theta, NPARAM)) || (s = lalloc(&L->grad, NPARAM)) || (s = lalloc(&L->m1, NPARAM)) ||
      (s = lalloc(&L->m2, NPARAM)) || (s = lalloc(&L->wb, NPARAM)) || (s = lalloc(&L->wt, NPARAM)) ||
      (s = lalloc(&L->bbits, 8)) || (s = lalloc(&L->efeat, Np * SL * 4)) || (s = lalloc(&L->X, Np * WD)) ||
      (s = lalloc(&L->Aed, Np * SL * WD)) || (s = lalloc(&L->H, (NL + 1) * Np * WD)) ||
      (s = lalloc(&L->Msg, NL * Np * SL * WD)) || (s = lalloc(&L->Agg, NL * Np * WD)) ||
      (s = lalloc(&L->Rr, NL * Np * WD)) || (s = lalloc(&L->P, (size_t)k * NB * N)) ||
      (s = lalloc(&L->G, Np * SL * 8)))
    return s;
  if (!im->LM) {
    const size_t es = im->prec == 32 ? sizeof(float) : sizeof(double);
    BES_HIP(hipMalloc(&im->LM, (size_t)(k + 1) * NB * N * es));
    BES_HIP(hipMemset(im->LM, 0, (size_t)(k + 1) * NB * N * es));
  }
  // static edge features in device cell order: outward normal, area / V^(2/3)
  const Geometry &g = c->geo;
  std::vector<float> ef(Np * SL * 4, 0.f);
  for (int i = 0; i < N; i++) {
    const int64_t e = c->perm[i];
    for (int m = 0; m < k; m++) {
      const int64_t f = g.cface[e * k + m];
      const double sg = g.owner[f] == e ? 1.0 : -1.0;
      for (int d = 0; d < nd; d++) ef[((size_t)i * SL + m) * 4 + d] = (float)(sg * g.n[f * nd + d]);
      ef[((size_t)i * SL + m) * 4 + nd] = (float)(g.area[f] / std::pow(g.vol[e], 2.0 / 3.0));
    }
  }
  BES_HIP(hipMemcpy(L->efeat, ef.data(), ef.size() * sizeof(float), hipMemcpyHostToDevice));
  std::vector<float> p0 = init_params(seed, nv);
  BES_HIP(hipMemcpy(L->theta, p0.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice));
  BES_HIP(hipMemset(L->m1, 0, NPARAM * sizeof(float)));
  BES_HIP(hipMemset(L->m2, 0, NPARAM * sizeof(float)));
  BES_HIP(hipMemset(L->grad, 0, NPARAM * sizeof(float)));
  BES_HIP(hipMemset(L->wt, 0, NPARAM * sizeof(uint16_t)));
  L->adam_t = 0;
  return lp_refresh(c, L);
}

static int lp_alloc_train(bes_ctx *c, Lp *L, int Q, int K) {
  if (L->train_ready && Q <= L->qcap && K <= L->kcap) return BES_OK;
  const size_t Np = (size_t)L->ntile * 32, n = (size_t)L->N * L->nv;
  void *old[] = {L->dH, L->dH2, L->dA, L->dnb, L->dMe, L->dlog, L->chain, L->tv1, L->tv2, L->gv, L->part, L->dsum,
                 L->gpart};
  for (void *p : old)
    if (p) hipFree(p);
  int s;
  if ((s = lalloc(&L->dH, Np * WD)) || (s = lalloc(&L->dH2, Np * WD)) || (s = lalloc(&L->dA, Np * WD)) ||
      (s = lalloc(&L->dnb, Np * SL * WD)) || (s = lalloc(&L->dMe, (size_t)L->k * L->nv * L->nv * L->N)) ||
      (s = lalloc(&L->dlog, Np * SL * WD)) || (s = lalloc(&L->chain, (size_t)Q * (K + 1) * n)) ||
      (s = lalloc(&L->tv1, n)) || (s = lalloc(&L->tv2, n)) || (s = lalloc(&L->gv, n)) ||
      (s = lalloc(&L->part, kDotBlocks)) || (s = lalloc(&L->dsum, 1)) ||
      (s = lalloc(&L->gpart, (size_t)kTrainWaves * NPARAM)))
    return s;
  BES_HIP(hipMemset(L->dH, 0, Np * WD * sizeof(float)));
  BES_HIP(hipMemset(L->gpart, 0, (size_t)kTrainWaves * NPARAM * sizeof(float)));
  L->qcap = Q;
  L->kcap = K;
  L->train_ready = true;
  return BES_OK;
}

static Net net_of(Lp *L) { return Net{L->wb, L->wt, L->theta}; }

// features + network forward + decode: M for the current system (Impl::A, Impl::b)
template <typename TM>
static int lp_build_t(bes_ctx *c, Impl *im) {
  Lp *L = lp_of(im);
  const int N = c->N, bs = 256, nbk = (L->ntile + 3) / 4;
  const size_t Np = (size_t)L->ntile * 32;
  ProfScope ps(c, kProfLearned, 0.0);
  hipLaunchKernelGGL(k_fill_f, dim3(1), dim3(64), 0, c->stream, (float *)L->bbits, 0.f, (size_t)16);
  if (c->nd == 2) {
    hipLaunchKernelGGL(k_bmax<2>, dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, im->b, N, L->bbits);
    hipLaunchKernelGGL((k_features<2, TM>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const TM *)im->A, im->b,
                       L->bbits, c->dm, L->efeat, L->X, L->Aed, (TM *)im->LM);
    hipLaunchKernelGGL((k_neumann<2, TM>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const TM *)im->A,
                       (const TM *)im->LM, c->dm, L->P);
  } else {
    hipLaunchKernelGGL(k_bmax<3>, dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, im->b, N, L->bbits);
    hipLaunchKernelGGL((k_features<3, TM>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const TM *)im->A, im->b,
                       L->bbits, c->dm, L->efeat, L->X, L->Aed, (TM *)im->LM);
    hipLaunchKernelGGL((k_neumann<3, TM>), dim3(grid_for(N, bs)), dim3(bs), 0, c->stream, (const TM *)im->A,
                       (const TM *)im->LM, c->dm, L->P);
  }
  const Net net = net_of(L);
  hipLaunchKernelGGL(k_encode, dim3(nbk), dim3(256), 0, c->stream, L->X, net, N, L->H);
  for (int l = 0; l < NL; l++)
    hipLaunchKernelGGL(k_layer, dim3(nbk), dim3(256), 0, c->stream, l, L->H + l * Np * WD, L->Aed, c->dm.nbr, c->k,
                       net, N, L->Msg + l * Np * SL * WD, L->Agg + l * Np * WD, L->Rr + l * Np * WD,
                       L->H + (l + 1) * Np * WD);
  if (c->nd == 2)
    hipLaunchKernelGGL((k_decode<2, TM>), dim3(nbk), dim3(256), 0, c->stream, L->H + NL * Np * WD, c->dm, net,
                       L->P, L->G, (TM *)im->LM);
  else
    hipLaunchKernelGGL((k_decode<3, TM>), dim3(nbk), dim3(256), 0, c->stream, L->H + NL * Np * WD, c->dm, net,
                       L->P, L->G, (TM *)im->LM);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

int lp_build(bes_ctx *c, Impl *im) {
  if (!im->lp) { set_error("learned preconditioner not initialised (bes_lp_init)"); return BES_ERR_ARG; }
  if (im->prec == 64) return lp_build_t<double>(c, im);
  int s = lp_build_t<float>(c, im);
  if (s) return s;
  const size_t nm = (size_t)(c->k + 1) * c->nv * c->nv * c->N;
  if (!im->LMh) BES_HIP(hipMalloc((void **)&im->LMh, nm * sizeof(__half)));
  hipLaunchKernelGGL(k_to_half, dim3(2048), dim3(256), 0, c->stream, (const float *)im->LM, im->LMh, nm);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

static double read_sumsq(bes_ctx *c, Lp *L, const float *x, size_t n, int *st) {
  hipLaunchKernelGGL(k_sumsq, dim3(kDotBlocks), dim3(256), 0, c->stream, x, n, L->part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->stream, L->part, kDotBlocks, L->dsum);
  double v = 0;
  hipError_t e = hipMemcpyAsync(&v, L->dsum, sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  *st = e == hipSuccess ? BES_OK : hip_fail(e, "sum of squares");
  return v;
}

template <typename TM>
static void spmv_f(bes_ctx *c, const TM *A, const float *x, float *y) {
  const int bs = 256;
  if (c->nd == 2) hipLaunchKernelGGL((k_spmv<2, TM, float>), dim3(grid_for(c->N, bs)), dim3(bs), 0, c->stream, A, c->dm, x, y, 0);
  else hipLaunchKernelGGL((k_spmv<3, TM, float>), dim3(grid_for(c->N, bs)), dim3(bs), 0, c->stream, A, c->dm, x, y, 0);
}
template <typename TM>
static void spmv_t_f(bes_ctx *c, Impl *im, const TM *A, const float *x, float *y) {
  const int bs = 256;
  if (c->nd == 2)
    hipLaunchKernelGGL((k_spmv_t<2, TM, float>), dim3(grid_for(c->N, bs)), dim3(bs), 0, c->stream, A, c->dm, im->fd.mrev, x, y);
  else
    hipLaunchKernelGGL((k_spmv_t<3, TM, float>), dim3(grid_for(c->N, bs)), dim3(bs), 0, c->stream, A, c->dm, im->fd.mrev, x, y);
}

// one forward + loss + backward on the current system; probes already in
// chain[q][0]; leaves the gradient in L->grad and returns the loss
template <typename TM>
static int lp_loss_grad_t(bes_ctx *c, Impl *im, int Q, int K, double *loss) {
  Lp *L = lp_of(im);
  const int N = c->N, nv = c->nv, NB = nv * nv;
  const size_t n = (size_t)N * nv, Np = (size_t)L->ntile * 32;
  int s;
  if ((s = lp_build_t<TM>(c, im))) return s;
  const TM *A = (const TM *)im->A, *M = (const TM *)im->LM;
  ProfScope ps(c, kProfLearned, 0.0);
  hipLaunchKernelGGL(k_fill_f, dim3(1024), dim3(256), 0, c->stream, L->dMe, 0.f, (size_t)c->k * NB * N);
  double Lsum = 0;
  for (int q = 0; q < Q; q++) {
    float *ch = L->chain + (size_t)q * (K + 1) * n;
    for (int kk = 0; kk < K; kk++) {
      spmv_f<TM>(c, M, ch + kk * n, L->tv1);
      spmv_f<TM>(c, A, L->tv1, L->tv2);
      hipLaunchKernelGGL(k_sub, dim3(1024), dim3(256), 0, c->stream, ch + kk * n, L->tv2, ch + (kk + 1) * n, n);
    }
    const double vv = read_sumsq(c, L, ch, n, &s);
    if (s) return s;
    const double rr = read_sumsq(c, L, ch + K * n, n, &s);
    if (s) return s;
    if (!(vv > 0)) { set_error("zero probe vector"); return BES_ERR_ARG; }
    Lsum += rr / vv / Q;
    // adjoint: g_K = 2 r_K / (Q vv); dMe -= (A^T g_{k+1}) (x) r_k; g_k = g_{k+1} - M^T A^T g_{k+1}
    hipLaunchKernelGGL(k_scale_f, dim3(1024), dim3(256), 0, c->stream, ch + K * n, (float)(2.0 / (Q * vv)), L->gv, n);
    for (int kk = K - 1; kk >= 0; kk--) {
      spmv_t_f<TM>(c, im, A, L->gv, L->tv1);
      if (c->nd == 2) hipLaunchKernelGGL(k_dme<2>, dim3(grid_for(N, 256)), dim3(256), 0, c->stream, L->tv1, ch + kk * n, c->dm, L->dMe);
      else hipLaunchKernelGGL(k_dme<3>, dim3(grid_for(N, 256)), dim3(256), 0, c->stream, L->tv1, ch + kk * n, c->dm, L->dMe);
      if (kk > 0) {
        spmv_t_f<TM>(c, im, M, L->tv1, L->tv2);
        hipLaunchKernelGGL(k_sub, dim3(1024), dim3(256), 0, c->stream, L->gv, L->tv2, L->gv, n);
      }
    }
  }
  *loss = Lsum;
  // back through the decoder, the layers and the encoder
  const Net net = net_of(L);
  if (c->nd == 2)
    hipLaunchKernelGGL(k_dlogit<2>, dim3(grid_for(N * SL, 256)), dim3(256), 0, c->stream, L->dMe, L->P, L->G, c->dm, L->dlog);
  else
    hipLaunchKernelGGL(k_dlogit<3>, dim3(grid_for(N * SL, 256)), dim3(256), 0, c->stream, L->dMe, L->P, L->G, c->dm, L->dlog);
  // rows of gpart beyond the launched waves were zeroed at allocation
  const int nblk = std::min(kTrainBlocks, (L->ntile + 3) / 4);
  hipLaunchKernelGGL(k_decode_bwd, dim3(nblk), dim3(256), 0, c->stream, L->dlog, L->H + NL * Np * WD, c->dm.nbr,
                     c->k, net, N, L->ntile, L->dA, L->dnb, L->gpart);
  hipLaunchKernelGGL(k_gather_rev, dim3(grid_for(N * WD, 256)), dim3(256), 0, c->stream, L->dA, L->dnb, c->dm.nbr,
                     im->fd.mrev, c->k, N, L->dH);
  for (int l = NL - 1; l >= 0; l--) {
    hipLaunchKernelGGL(k_layer_bwd, dim3(nblk), dim3(256), 0, c->stream, l, L->dH, L->H + l * Np * WD,
                       L->Agg + l * Np * WD, L->Rr + l * Np * WD, L->Msg + l * Np * SL * WD, L->Aed, c->dm.nbr, c->k,
                       net, N, L->ntile, L->dH2, L->dnb, L->gpart);
    hipLaunchKernelGGL(k_gather_rev, dim3(grid_for(N * WD, 256)), dim3(256), 0, c->stream, L->dH2, L->dnb, c->dm.nbr,
                       im->fd.mrev, c->k, N, L->dH);
  }
  hipLaunchKernelGGL(k_encode_bwd, dim3(nblk), dim3(256), 0, c->stream, L->dH, L->H, L->X, N, L->ntile, L->gpart);
  hipLaunchKernelGGL(k_grad_reduce, dim3(grid_for(NPARAM, 256)), dim3(256), 0, c->stream, L->gpart, kTrainWaves,
                     L->grad);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

static int lp_loss_grad(bes_ctx *c, Impl *im, int Q, int K, double *loss) {
  return im->prec == 64 ? lp_loss_grad_t<double>(c, im, Q, K, loss) : lp_loss_grad_t<float>(c, im, Q, K, loss);
}

// probes: q = 0 the right-hand side, q >= 1 seeded Gaussian vectors
static int lp_default_probes(bes_ctx *c, Impl *im, int Q, int K, uint64_t seed) {
  Lp *L = lp_of(im);
  const size_t n = (size_t)c->N * c->nv;
  hipLaunchKernelGGL(k_from_double, dim3(1024), dim3(256), 0, c->stream, im->b, L->chain, n);
  for (int q = 1; q < Q; q++)
    hipLaunchKernelGGL(k_randn, dim3(1024), dim3(256), 0, c->stream, (unsigned long long)(seed * 1000003ull + q),
                       L->chain + (size_t)q * (K + 1) * n, n);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

}  // namespace lp
}  // namespace bes

using namespace bes;
using namespace bes::lp;

static Impl *impl_get(bes_ctx *c) { return c ? (Impl *)c->impl : nullptr; }

extern "C" {

int bes_lp_param_count(void) { return NPARAM; }

int bes_lp_init(bes_ctx *c, unsigned long long seed) {
  Impl *im = impl_get(c);
  if (!im) { set_error("call bes_implicit_setup first"); return BES_ERR_ARG; }
  im->lp_current = false;
  return lp_alloc(c, im, seed);
}

int bes_lp_set_rebuild_interval(bes_ctx *c, int n) {
  Impl *im = impl_get(c);
  if (!im || n < 1) { set_error("call bes_implicit_setup first / interval >= 1"); return BES_ERR_ARG; }
  im->lp_rebuild = n;
  im->lp_current = false;
  return BES_OK;
}

long bes_lp_builds(bes_ctx *c) {
  Impl *im = impl_get(c);
  return im ? im->lp_builds : 0;
}

int bes_lp_set_params(bes_ctx *c, const float *params, int n) {
  Impl *im = impl_get(c);
  if (!im || !im->lp || !params || n != NPARAM) { set_error("bes_lp_init first / parameter count"); return BES_ERR_ARG; }
  Lp *L = lp_of(im);
  im->lp_current = false;
  im->factored = false;
  BES_HIP(hipMemcpy(L->theta, params, NPARAM * sizeof(float), hipMemcpyHostToDevice));
  BES_HIP(hipMemset(L->m1, 0, NPARAM * sizeof(float)));
  BES_HIP(hipMemset(L->m2, 0, NPARAM * sizeof(float)));
  L->adam_t = 0;
  int s = lp_refresh(c, L);
  if (s) return s;
  BES_HIP(hipStreamSynchronize(c->stream));
  return BES_OK;
}

int bes_lp_get_params(bes_ctx *c, float *params, int n) {
  Impl *im = impl_get(c);
  if (!im || !im->lp || !params || n != NPARAM) { set_error("bes_lp_init first / parameter count"); return BES_ERR_ARG; }
  BES_HIP(hipStreamSynchronize(c->stream));
  BES_HIP(hipMemcpy(params, lp_of(im)->theta, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return BES_OK;
}

int bes_lp_operator(bes_ctx *c, double *Md, double *Me, double *G) {
  Impl *im = impl_get(c);
  if (!im || !im->lp || !im->assembled) { set_error("assemble and bes_lp_init first"); return BES_ERR_ARG; }
  int s = lp_build(c, im);
  if (s) return s;
  Lp *L = lp_of(im);
  const int N = c->N, k = c->k, nv = c->nv, NB = nv * nv;
  const size_t es = im->prec == 32 ? sizeof(float) : sizeof(double);
  std::vector<char> raw((size_t)(k + 1) * NB * N * es);
  std::vector<float> gb((size_t)N * SL * 8);
  BES_HIP(hipMemcpyAsync(raw.data(), im->LM, raw.size(), hipMemcpyDeviceToHost, c->stream));
  BES_HIP(hipMemcpyAsync(gb.data(), L->G, gb.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  auto val = [&](size_t q) -> double {
    return im->prec == 32 ? (double)((const float *)raw.data())[q] : ((const double *)raw.data())[q];
  };
  for (int i = 0; i < N; i++) {
    const int64_t e = c->perm[i];
    for (int q = 0; q < NB; q++) {
      if (Md) Md[e * NB + q] = val((size_t)q * N + i);
      for (int m = 0; m < k; m++)
        if (Me) Me[(e * k + m) * NB + q] = val(((size_t)(1 + m) * NB + q) * N + i);
    }
    if (G)
      for (int m = 0; m < k; m++)
        for (int r = 0; r < nv; r++) G[(e * k + m) * nv + r] = gb[((size_t)i * SL + m) * 8 + r];
  }
  return BES_OK;
}

// loss and gradient for caller-given probe vectors [nprobe][ncell][nv] (reference order)
int bes_lp_loss_grad(bes_ctx *c, const double *probes, int nprobe, int nrich, double *loss, float *grad) {
  Impl *im = impl_get(c);
  if (!im || !im->lp || !im->assembled || !probes || nprobe < 1 || nprobe > 16 || nrich < 1 || nrich > 8) {
    set_error("assemble and bes_lp_init first / probe count 1-16, Richardson steps 1-8");
    return BES_ERR_ARG;
  }
  Lp *L = lp_of(im);
  int s = lp_alloc_train(c, L, nprobe, nrich);
  if (s) return s;
  const size_t n = (size_t)c->N * c->nv;
  std::vector<double> tmp(n);
  std::vector<float> f(n);
  for (int q = 0; q < nprobe; q++) {
    for (int i = 0; i < c->N; i++) {
      const int64_t e = c->perm[i];
      for (int v = 0; v < c->nv; v++) f[(size_t)i * c->nv + v] = (float)probes[((size_t)q * c->N + e) * c->nv + v];
    }
    BES_HIP(hipMemcpyAsync(L->chain + (size_t)q * (nrich + 1) * n, f.data(), n * sizeof(float), hipMemcpyHostToDevice,
                           c->stream));
    BES_HIP(hipStreamSynchronize(c->stream));
  }
  double Lv = 0;
  if ((s = lp_loss_grad(c, im, nprobe, nrich, &Lv))) return s;
  if (loss) *loss = Lv;
  if (grad) {
    BES_HIP(hipMemcpyAsync(grad, L->grad, NPARAM * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    BES_HIP(hipStreamSynchronize(c->stream));
  }
  return prof_collect(c);
}

// Adam steps on the currently assembled system; loss_hist [steps] may be NULL
int bes_lp_train(bes_ctx *c, int steps, double lr, int nprobe, int nrich, unsigned long long seed, double *loss_hist) {
  Impl *im = impl_get(c);
  if (!im || !im->lp || !im->assembled || steps < 0 || !(lr > 0) || nprobe < 1 || nprobe > 16 || nrich < 1 ||
      nrich > 8) {
    set_error("assemble and bes_lp_init first / steps >= 0, lr > 0, probes 1-16, Richardson steps 1-8");
    return BES_ERR_ARG;
  }
  Lp *L = lp_of(im);
  int s = lp_alloc_train(c, L, nprobe, nrich);
  if (s) return s;
  for (int it = 0; it < steps; it++) {
    if ((s = lp_default_probes(c, im, nprobe, nrich, seed + (uint64_t)it))) return s;
    double Lv = 0;
    if ((s = lp_loss_grad(c, im, nprobe, nrich, &Lv))) return s;
    if (loss_hist) loss_hist[it] = Lv;
    L->adam_t++;
    const float c1 = 1.f - std::pow(0.9f, (float)L->adam_t), c2 = 1.f - std::pow(0.999f, (float)L->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(NPARAM, 256)), dim3(256), 0, c->stream, L->theta, L->grad, L->m1, L->m2,
                       (float)lr, c1, c2, L->wb, L->wt);
    BES_HIP(hipGetLastError());
  }
  im->factored = false;   // the operator changed: rebuild before the next solve
  im->lp_current = false;
  BES_HIP(hipStreamSynchronize(c->stream));
  return prof_collect(c);
}

int bes_lp_set_online(bes_ctx *c, int steps, double lr, int nprobe, int nrich) {
  Impl *im = impl_get(c);
  if (!im || !im->lp || steps < 0 || !(lr > 0) || nprobe < 1 || nprobe > 16 || nrich < 1 || nrich > 8) {
    set_error("bes_lp_init first / steps >= 0, lr > 0, probes 1-16, Richardson steps 1-8");
    return BES_ERR_ARG;
  }
  im->lp_online_steps = steps;
  im->lp_online_lr = lr;
  im->lp_online_nprobe = nprobe;
  im->lp_online_nrich = nrich;
  return BES_OK;
}

double bes_lp_online_loss(bes_ctx *c) {
  Impl *im = impl_get(c);
  return im ? im->lp_online_loss : 0.0;
}

int bes_lp_fallbacks(bes_ctx *c) {
  Impl *im = impl_get(c);
  return im ? im->lp_fallbacks : 0;
}

}  // extern "C"
