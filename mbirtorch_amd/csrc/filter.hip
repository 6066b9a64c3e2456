// Direct-reconstruction row filter and sinogram layout transposes (gfx950).
//
// The reference convolves every detector row with the 2C-1 ramp taps by
// FFT (tomography_utils.py:29-78).  Here the 'valid' convolution is computed
// directly: a workgroup stages four detector rows of one view (pre-weighted
// by the optional FDK cosine map) and the taps in LDS, and each thread owns
// one channel of the four rows:
//     out[v, r, c] = sum_k in[v, r, k] * row_weight[r, k] * taps[c - k + C - 1].
// The output can be written channel-major, ready for the back projector.
#include "common.h"

namespace mbir {

constexpr int kFiltRows = 4;
constexpr int kFiltThreads = 256;

__global__ __launch_bounds__(kFiltThreads) void row_filter_kernel(
    const float* __restrict__ in, const float* __restrict__ taps,
    const float* __restrict__ rw, int V, int R, int C, int out_vcr, float* __restrict__ out) {
  extern __shared__ float sm[];
  float* s_taps = sm;                   // 2C - 1
  float* s_rows = sm + (2 * C - 1);     // kFiltRows * C
  const int v = blockIdx.x, r0 = blockIdx.y * kFiltRows;
  for (int i = threadIdx.x; i < 2 * C - 1; i += kFiltThreads) s_taps[i] = taps[i];
  for (int i = threadIdx.x; i < kFiltRows * C; i += kFiltThreads) {
    const int j = i / C, k = i - j * C;
    const int r = r0 + j;
    float val = 0.0f;
    if (r < R) {
      val = in[((size_t)v * R + r) * C + k];
      if (rw) val = val * rw[(size_t)r * C + k];
    }
    s_rows[i] = val;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += kFiltThreads) {
    float acc[kFiltRows];
#pragma unroll
    for (int j = 0; j < kFiltRows; ++j) acc[j] = 0.0f;
    const float* t = s_taps + c + C - 1;  // t[-k] = taps[c - k + C - 1]
    for (int k = 0; k < C; ++k) {
      const float h = t[-k];
#pragma unroll
      for (int j = 0; j < kFiltRows; ++j) acc[j] = fmaf(s_rows[j * C + k], h, acc[j]);
    }
#pragma unroll
    for (int j = 0; j < kFiltRows; ++j) {
      const int r = r0 + j;
      if (r >= R) continue;
      if (out_vcr) out[((size_t)v * C + c) * R + r] = acc[j];
      else out[((size_t)v * R + r) * C + c] = acc[j];
    }
  }
}

// 32x32 tiled transpose of each view's (A, B) matrix into (B, A).
__global__ void transpose_kernel(const float* __restrict__ in, int V, int A, int B,
                                 float* __restrict__ out) {
  __shared__ float tile[32][33];
  const int v = blockIdx.z;
  const int a0 = blockIdx.y * 32, b0 = blockIdx.x * 32;
  const float* src = in + (size_t)v * A * B;
  float* dst = out + (size_t)v * A * B;
  for (int j = threadIdx.y; j < 32; j += blockDim.y) {
    const int a = a0 + j, b = b0 + threadIdx.x;
    if (a < A && b < B) tile[j][threadIdx.x] = src[(size_t)a * B + b];
  }
  __syncthreads();
  for (int j = threadIdx.y; j < 32; j += blockDim.y) {
    const int b = b0 + j, a = a0 + threadIdx.x;
    if (a < A && b < B) dst[(size_t)b * A + a] = tile[threadIdx.x][j];
  }
}

}  // namespace mbir

using namespace mbir;

extern "C" {

int mbir_row_filter(const float* in, const float* taps, const float* rw, int32_t V, int32_t R,
                    int32_t C, int32_t out_layout, float* out, mbir_stream_t stream) {
  if (!in || !taps || !out) return MBIR_ERR_ARG;
  if (V <= 0 || R <= 0 || C <= 0) return MBIR_ERR_SHAPE;
  if (out_layout != MBIR_LAYOUT_VRC && out_layout != MBIR_LAYOUT_VCR) return MBIR_ERR_ARG;
  const size_t lds = (size_t)(2 * C - 1 + kFiltRows * C) * sizeof(float);
  if (lds > 60 * 1024) return MBIR_ERR_SHAPE;
  dim3 grid(V, (R + kFiltRows - 1) / kFiltRows);
  row_filter_kernel<<<grid, kFiltThreads, lds, (hipStream_t)stream>>>(
      in, taps, rw, V, R, C, out_layout == MBIR_LAYOUT_VCR, out);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int mbir_transpose_sino(const float* in, int32_t V, int32_t R, int32_t C, int32_t to_vcr,
                        float* out, mbir_stream_t stream) {
  if (!in || !out || in == out) return MBIR_ERR_ARG;
  if (V <= 0 || R <= 0 || C <= 0) return MBIR_ERR_SHAPE;
  // to_vcr: each view is (R, C) -> (C, R); otherwise (C, R) -> (R, C).
  const int A = to_vcr ? R : C, B = to_vcr ? C : R;
  dim3 grid((B + 31) / 32, (A + 31) / 32, V);
  transpose_kernel<<<grid, dim3(32, 8), 0, (hipStream_t)stream>>>(in, V, A, B, out);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

}  // extern "C"
