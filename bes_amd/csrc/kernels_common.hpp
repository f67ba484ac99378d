// Launch-shape helpers shared by the BES kernels (gfx950: wave64, 8 XCDs).
#pragma once

#include <hip/hip_runtime.h>

namespace bes {

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  Remap the hardware block id so that XCD x processes one
// contiguous range of blocks: Morton-ordered neighbours then share an L2.
// The map is a bijection on [0, gridDim.x) for any grid size.
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int full = nb >> 3, rem = nb & 7, x = b & 7, q = b >> 3;
  const int start = x * full + (x < rem ? x : rem);
  return start + q;
}

__device__ __forceinline__ int cell_index() { return xcd_block() * blockDim.x + threadIdx.x; }

// Quotients a / b for one divisor b, bitwise equal to the correctly rounded
// division, with one division in all.  y = RN(1/b) is computed once; then
// q = RN(a*y) is within an ulp of a/b, its remainder a - q*b is exact in one
// FMA, and RN(q + r*y) is the correctly rounded quotient (Markstein's theorem,
// which needs no underflow or overflow on the way: hence the ranges, outside
// of which the plain division is taken).
struct RcpDiv {
  double b, y;
  bool ok;
  __host__ __device__ explicit RcpDiv(double b_)
      : b(b_), y(1.0 / b_), ok(__builtin_fabs(b_) >= 0x1p-500 && __builtin_fabs(b_) <= 0x1p+500) {}
  __host__ __device__ double operator()(double a) const {
    const double aa = __builtin_fabs(a);
    if (ok && aa >= 0x1p-400 && aa <= 0x1p+400) {
      const double q = a * y;
      return __builtin_fma(__builtin_fma(-q, b, a), y, q);
    }
    return a / b;
  }
};

// one value per condition of a batch (at most 64), passed to a kernel by value
struct CoefB {
  double c[64];
};

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block reductions (blockDim.x a multiple of 64, at most 1024).  Every thread
// of the block must call them; the result is valid in thread 0.
__device__ __forceinline__ double block_max(double v) {
  __shared__ double s[16];
  v = wave_max(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) s[w] = v;
  __syncthreads();
  double r = -__builtin_inf();
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) r = fmax(r, s[i]);
  return r;
}
__device__ __forceinline__ double block_min(double v) {
  __shared__ double s[16];
  v = wave_min(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) s[w] = v;
  __syncthreads();
  double r = __builtin_inf();
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) r = fmin(r, s[i]);
  return r;
}

// Per-condition block reductions for batched launches: thread t of a block
// works on condition t % B (B a power of two <= 64, blockDim.x a multiple of
// 64).  Lanes of a wave with the same condition are combined by xor-shuffles
// with offsets >= B, then the waves through LDS; thread t < B holds the
// result for condition t.  Every thread of the block must call them.
// Linf monitors: a launch's blocks add their maxima with atomicMax into
// kLinfSlots copies of the monitor words (block b into copy b mod kLinfSlots),
// not one, so thousands of blocks do not queue on a few addresses; a reduce
// over the copies follows (max is exact, so the order does not matter).
constexpr int kLinfSlots = 64;
__device__ __forceinline__ unsigned long long *linf_slot_of(unsigned long long *bits, int nq, int blk) {
  return bits + (size_t)(blk & (kLinfSlots - 1)) * nq;
}
__device__ __forceinline__ unsigned long long *linf_slot(unsigned long long *bits, int nq) {
  return linf_slot_of(bits, nq, blockIdx.x);
}
// out[q] = max over the copies of bits[s*nq + q] (out may be bits: copy 0)
__device__ __forceinline__ unsigned long long linf_slots_max(const unsigned long long *bits, int nq, int q) {
  unsigned long long m = 0;
  for (int s = 0; s < kLinfSlots; s++) {
    const unsigned long long x = bits[(size_t)s * nq + q];
    m = x > m ? x : m;
  }
  return m;
}

__device__ __forceinline__ double block_max_cond(double v, int B) {
  __shared__ double s[16][64];
  for (int o = 32; o >= B; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l < B) s[w][l] = v;
  __syncthreads();
  double r = -__builtin_inf();
  if ((int)threadIdx.x < B)
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) r = fmax(r, s[q][threadIdx.x]);
  return r;
}
__device__ __forceinline__ double block_min_cond(double v, int B) {
  __shared__ double s[16][64];
  for (int o = 32; o >= B; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l < B) s[w][l] = v;
  __syncthreads();
  double r = __builtin_inf();
  if ((int)threadIdx.x < B)
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) r = fmin(r, s[q][threadIdx.x]);
  return r;
}

}  // namespace bes
