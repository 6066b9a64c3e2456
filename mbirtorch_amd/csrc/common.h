// Shared device helpers for the MI355X (gfx950) projector / VCD kernels.
//
// All kernels are written for 64-wide wavefronts.  Geometry arithmetic is
// compiled with -ffp-contract=off so every add/mul/div rounds on its own, in
// the order the reference's eager float32 chain uses (parallel_beam.py:22-53,
// cone_beam.py:59-158); accumulations call fmaf() explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mbir_hip.h"

#define MBIR_WAVE 64

#define MBIR_CHECK_LAUNCH()                                   \
  do {                                                        \
    hipError_t e_ = hipGetLastError();                        \
    if (e_ != hipSuccess) return MBIR_ERR_LAUNCH;             \
  } while (0)

namespace mbir {

// Per-view constants of the parallel chain, computed on the host in float32
// exactly as the reference's eager torch ops do (cos/sin of the float32
// angle, footprint, W, weight scale), then uploaded once per model.
struct ParView {
  float cosv, sinv;   // cos/sin of the view angle
  float W;            // projected voxel width in channel units
  float ws;           // weight scale = dv*dvr / footprint
  float halfw;        // (W + 1) / 2
  float lmax;         // min(W, 1)
  float pad0, pad1;
};

// Sinogram storage: float32, or bfloat16 kept as its 16 raw bits (the high
// half of the float32 pattern).  Loads widen exactly; stores round to
// nearest even (finite values).
typedef uint16_t bf16_bits;

__device__ __forceinline__ float bf16_to_f32(bf16_bits h) {
  return __uint_as_float((uint32_t)h << 16);
}

__device__ __forceinline__ bf16_bits f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16_bits)(u >> 16);
}

__device__ __forceinline__ float ld_s(const float* p) { return *p; }
__device__ __forceinline__ float ld_s(const bf16_bits* p) { return bf16_to_f32(*p); }
__device__ __forceinline__ void st_s(float* p, float v) { *p = v; }
__device__ __forceinline__ void st_s(bf16_bits* p, float v) { *p = f32_to_bf16(v); }

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, MBIR_WAVE);
  return v;
}

__device__ __forceinline__ float wave_sum_f32(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, MBIR_WAVE);
  return v;
}

// Block-wide FP64 sum of two values; result valid in thread 0.  Requires
// blockDim.x a multiple of 64 and <= 1024.  Deterministic: fixed tree.
__device__ __forceinline__ void block_sum2_f64(double& a, double& b) {
  __shared__ double red_a[16], red_b[16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  a = wave_sum_f64(a);
  b = wave_sum_f64(b);
  __syncthreads();
  if (lane == 0) { red_a[wid] = a; red_b[wid] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0;
    const int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; ++i) { sa += red_a[i]; sb += red_b[i]; }
    a = sa; b = sb;
  }
}

// XCD-aware block order.  Workgroups are dispatched round-robin over the
// MI355X's 8 XCDs (linear id mod 8), each with its own L2.  xcd_remap maps
// the linear dispatch id to a logical id so that XCD x runs one contiguous
// range of logical ids -- neighbouring tiles (which read the same sinogram
// rows or voxel cylinders) then share an L2.  A bijection on [0, n).
constexpr int kNumXcd = 8;

__device__ __forceinline__ int xcd_remap(int id, int n) {
  const int per = n / kNumXcd, rem = n % kNumXcd;
  const int xcd = id % kNumXcd, k = id / kNumXcd;
  return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

// Rank of this lane's bin b (b < 0: no entry) among the lanes below it with
// the same bin, the size of its group, and whether it is the group's lowest
// lane.  One ballot per distinct bin in the wavefront.
__device__ __forceinline__ void wave_bin_rank(int b, int& rank, int& gsize, bool& leader) {
  const int lane = threadIdx.x & (MBIR_WAVE - 1);
  unsigned long long todo = __ballot(b >= 0);
  rank = 0;
  gsize = 0;
  leader = false;
  while (todo) {
    const int l = __ffsll(todo) - 1;
    const int bl = __shfl(b, l, MBIR_WAVE);
    const unsigned long long m = __ballot(b == bl);
    if (b == bl) {
      rank = __popcll(m & ((1ull << lane) - 1ull));
      gsize = __popcll(m);
      leader = lane == l;
    }
    todo &= ~m;
  }
}

}  // namespace mbir
