// Parallel-beam projectors for MI355X (gfx950).
//
// Geometry (one (view, pixel) pair): the reference's float32 chain
// parallel_beam.py:22-53 -- pixel row/col, x = cos*x~ - sin*y~,
// n_p = (x + offset)/delta_channel + (C-1)/2 -- followed by the shared
// trapezoid tap rule horizontal_fan.py:21-24,51-71:
//     A(n) = ws * min(max((W+1)/2 - |n_p - n|, 0), min(1, W)),  0 off-detector.
// When every view has W <= 1 the nonzero taps are exactly floor(n_p) and
// floor(n_p)+1 (|n - n_p| < (W+1)/2 <= 1), a subset of the reference's
// round(n_p) +- psf taps with identical weights, so two taps are evaluated;
// otherwise the reference's 2*psf+1 taps around round(n_p) are.  All three
// kernels below evaluate n_p with one function (par_np), so a (view, pixel,
// channel) weight is bit-identical in the forward and back projectors.
//
// Back projector (owner computes, atomic-free, deterministic): a workgroup
// owns a set of voxel cylinders and a view range.  Geometry for
// (view batch x pixels) is computed once into LDS; lanes of a wavefront then
// map to (pixel slot, detector row) so each tap reads a contiguous run of the
// channel-major sinogram row e[v, n, 0:R].  Small pixel sets split the view
// range over workgroups and leave partial sums reduced in a fixed order.
//
// Forward projector, two gather forms (atomic-free, deterministic):
//  * ray: for pixel lists in which every image row is one run of
//    consecutive columns (the full region of reconstruction), each thread
//    owns one detector channel of one view and a band of detector rows, walks
//    the image rows (or columns, whichever the view's projection is steeper
//    in), evaluates the 3-5 candidate pixels of the channel exactly, and
//    accumulates in registers;
//  * binned: for arbitrary pixel lists (the VCD subsets), a workgroup owns one
//    view and a block of channels; chunks of the list are counting-sorted in
//    LDS by first tap channel, each bin ordered by list position, and each
//    unit walks the contiguous run of sorted pixels whose taps reach its
//    channels, as the ray form walks image rows.  Its epilogue writes the
//    projection, or -- for the VCD
//    subset step -- reduces it against the error sinogram (line search, FP64)
//    or subtracts it from the error sinogram, so the subset direction's
//    projection never goes to HBM.
#include "common.h"

namespace mbir {

constexpr int kBlock = 256;
constexpr int kMaxTaps = 9;  // psf_radius <= 4

// Channel coordinate n_p of pixel (row, col) at one view.
__device__ __forceinline__ float par_np(int row, int col, const ParView& vw,
                                        const mbir_par_geom& g) {
  const float y_t = g.delta_voxel_row * ((float)row - g.row_center);
  const float x_t = g.delta_voxel * ((float)col - g.col_center);
  const float x = vw.cosv * x_t - vw.sinv * y_t;
  float t = x + g.det_channel_offset;
  if (!g.unit_channel) t = t / g.delta_det_channel;
  return t + g.det_center;
}

template <int NT>
__device__ __forceinline__ int par_base(float n_p, const mbir_par_geom& g) {
  return NT == 2 ? (int)floorf(n_p) : (int)rintf(n_p) - g.psf_radius;
}

__device__ __forceinline__ float par_weight(float n_p, int n, const ParView& vw) {
  const float A = fmaxf(vw.halfw - fabsf(n_p - (float)n), 0.0f);
  return fminf(A, vw.lmax) * vw.ws;
}

// Geometry for one (view, pixel): tap base index and NT weights (zero for
// taps off the detector).
template <int NT>
__device__ __forceinline__ void par_taps(int64_t pix, const ParView& vw,
                                         const mbir_par_geom& g, int& base,
                                         float (&a)[NT]) {
  // Pixel indices lie in [0, rows*cols) < 2^31: 32-bit division.
  const int pj = (int)pix;
  const int row = pj / g.num_cols;
  const int col = pj - row * g.num_cols;
  const float n_p = par_np(row, col, vw, g);
  base = par_base<NT>(n_p, g);
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    const int n = base + k;
    a[k] = (n >= 0 && n < g.num_channels) ? par_weight(n_p, n, vw) : 0.0f;
  }
}

// Weight of channel n for pixel (row, col) when n is one of the pixel's taps.
template <int NT>
__device__ __forceinline__ bool par_tap_of(int row, int col, int n, const ParView& vw,
                                           const mbir_par_geom& g, float& A) {
  const float n_p = par_np(row, col, vw, g);
  const int k = n - par_base<NT>(n_p, g);
  if (k < 0 || k >= NT) return false;
  A = par_weight(n_p, n, vw);
  return A != 0.0f;
}

// ---------------------------------------------------------------------------
// Back projection.
template <int RW, int NT, int CP, bool WEIGHTED, typename TS>
__global__ __launch_bounds__(kBlock) void par_back_kernel(
    const TS* __restrict__ sino, const TS* __restrict__ wts,
    const int64_t* __restrict__ pix, int n_pix, const ParView* __restrict__ views,
    mbir_par_geom g, int views_per_split, float* __restrict__ out) {
  constexpr int SLOTS = MBIR_WAVE / RW;
  constexpr int ACC = RW >= 4 ? 8 : 4;
  constexpr int PIX_WAVE = SLOTS * ACC;
  constexpr int PIX_WG = (kBlock / MBIR_WAVE) * PIX_WAVE;
  constexpr int VB = PIX_WG >= 1024 ? 1 : 1024 / PIX_WG;
  __shared__ int s_n0[VB * PIX_WG];
  __shared__ float s_a[NT * VB * PIX_WG];

  const int lane = threadIdx.x & (MBIR_WAVE - 1), wave = threadIdx.x / MBIR_WAVE;
  const int slot = lane / RW, rl = lane % RW;
  const int R = g.num_slices, C = g.num_channels;
  // With the view range split (small pixel sets), XCD-aware order: one XCD
  // runs consecutive pixel groups of one split, so the split's sinogram
  // lines are fetched into one L2 rather than all eight (measured 1.4-1.7x
  // at 1602-pixel subsets; a single split keeps the plain order).
  int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
  if (gridDim.y > 1) {
    const int nx = gridDim.x, ny = gridDim.y;
    const int l = xcd_remap(blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z), nx * ny * gridDim.z);
    bx = l % nx;
    by = (l / nx) % ny;
    bz = l / (nx * ny);
  }
  const int r = bz * RW + rl;
  const bool r_ok = r < R;
  const int p_base = bx * PIX_WG;
  const int v_begin = by * views_per_split;
  const int v_end = min(v_begin + views_per_split, g.num_views);

  float acc[ACC];
#pragma unroll
  for (int k = 0; k < ACC; ++k) acc[k] = 0.0f;

  for (int vb0 = v_begin; vb0 < v_end; vb0 += VB) {
    const int nvb = min(VB, v_end - vb0);
    __syncthreads();
    for (int q = threadIdx.x; q < VB * PIX_WG; q += kBlock) {
      const int j = q % PIX_WG, b = q / PIX_WG;
      const int pg = p_base + j;
      int base = 0;
      float a[NT];
      if (b < nvb && pg < n_pix) {
        par_taps<NT>(pix[pg], views[vb0 + b], g, base, a);
      } else {
#pragma unroll
        for (int k = 0; k < NT; ++k) a[k] = 0.0f;
      }
      s_n0[q] = base;
#pragma unroll
      for (int k = 0; k < NT; ++k) s_a[k * VB * PIX_WG + q] = CP == 2 ? a[k] * a[k] : a[k];
    }
    __syncthreads();
    if (r_ok) {
      for (int b = 0; b < nvb; ++b) {
        const size_t vofs = (size_t)(vb0 + b) * C * R + r;
        const TS* sv = sino + vofs;
        const TS* wv = WEIGHTED ? wts + vofs : nullptr;
        if constexpr (SLOTS == 1 && NT == 2) {
          // One pixel slot per wave: the ACC tap bases are wave-uniform.  When
          // they span at most ACC channels (neighbouring pixels of an image
          // row: adjacent bases differ by at most one since |cos| dv/ddc <= W
          // <= 1), the channel window [lo, hi+1] is loaded once and every tap
          // read from registers -- about 8|cos|+2 loads instead of 2*ACC.
          const int q0 = b * PIX_WG + wave * PIX_WAVE;
          int bk[ACC];
          int lo = s_n0[q0], hi = lo;
#pragma unroll
          for (int k = 0; k < ACC; ++k) {
            bk[k] = s_n0[q0 + k];
            lo = min(lo, bk[k]);
            hi = max(hi, bk[k]);
          }
          lo = __builtin_amdgcn_readfirstlane(lo);
          hi = __builtin_amdgcn_readfirstlane(hi);
          if (hi - lo < ACC) {
            const int len = hi - lo + 2;
            float win[ACC + 2];
#pragma unroll
            for (int m = 0; m < ACC + 2; ++m) {
              win[m] = 0.0f;
              if (m < len) {
                const int n = min(max(lo + m, 0), C - 1);
                float e = ld_s(sv + (size_t)n * R);
                if (WEIGHTED) e = e * ld_s(wv + (size_t)n * R);
                win[m] = e;
              }
            }
#pragma unroll
            for (int k = 0; k < ACC; ++k) {
              const int d = __builtin_amdgcn_readfirstlane(bk[k] - lo);  // 0 <= d < ACC
              acc[k] = fmaf(s_a[q0 + k], win[d], acc[k]);
              acc[k] = fmaf(s_a[VB * PIX_WG + q0 + k], win[d + 1], acc[k]);
            }
            continue;
          }
        }
#pragma unroll
        for (int k = 0; k < ACC; ++k) {
          const int q = b * PIX_WG + wave * PIX_WAVE + k * SLOTS + slot;
          const int base = s_n0[q];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const float a = s_a[t * VB * PIX_WG + q];
            const int n = min(max(base + t, 0), C - 1);
            float e = ld_s(sv + (size_t)n * R);
            if (WEIGHTED) e = e * ld_s(wv + (size_t)n * R);
            acc[k] = fmaf(a, e, acc[k]);
          }
        }
      }
    }
  }
  if (r_ok) {
#pragma unroll
    for (int k = 0; k < ACC; ++k) {
      const int pg = p_base + wave * PIX_WAVE + k * SLOTS + slot;
      if (pg < n_pix) out[((size_t)by * n_pix + pg) * R + r] = acc[k];
    }
  }
}

__global__ void reduce_splits_kernel(const float* __restrict__ parts, int ns,
                                     int64_t n, float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float s = parts[i];
    for (int k = 1; k < ns; ++k) s += parts[(size_t)k * n + i];
    out[i] = s;
  }
}

// ---------------------------------------------------------------------------
// Forward projectors.  Lane layout shared by both forms: a gather unit is LPC
// consecutive lanes of a wavefront, and lane rl of a unit owns detector rows
// rr = r0 + rl*VR .. rr+VR-1 of a row chunk of ROWS = LPC*VR rows.  A voxel
// cylinder segment is then one contiguous LPC*VR*4-byte load per unit, and a
// channel-major sinogram row segment (v, n, r0 : r0+ROWS) one contiguous
// store, as in the back projector.
template <int LPC, int VR>
struct FwdLanes {
  static constexpr int SPW = MBIR_WAVE / LPC;  // units per wavefront
  static constexpr int UNITS = kBlock / LPC;   // units per workgroup
  static constexpr int ROWS = LPC * VR;        // rows per chunk
};

template <int VR>
__device__ __forceinline__ void load_vr(const float* p, float (&x)[VR]) {
  if constexpr (VR % 4 == 0 && VR > 4) {
#pragma unroll
    for (int h = 0; h < VR; h += 4) {
      const float4 t = *reinterpret_cast<const float4*>(p + h);
      x[h] = t.x;
      x[h + 1] = t.y;
      x[h + 2] = t.z;
      x[h + 3] = t.w;
    }
  } else if constexpr (VR == 4) {
    const float4 t = *reinterpret_cast<const float4*>(p);
    x[0] = t.x;
    x[1] = t.y;
    x[2] = t.z;
    x[3] = t.w;
  } else {
#pragma unroll
    for (int q = 0; q < VR; ++q) x[q] = p[q];
  }
}

template <int VR>
__device__ __forceinline__ void store_vr(float* p, const float (&x)[VR]) {
  if constexpr (VR % 4 == 0) {
#pragma unroll
    for (int h = 0; h < VR; h += 4)
      *reinterpret_cast<float4*>(p + h) = make_float4(x[h], x[h + 1], x[h + 2], x[h + 3]);
  } else {
#pragma unroll
    for (int q = 0; q < VR; ++q) p[q] = x[q];
  }
}

// bfloat16 rows: two values per 32-bit word, element 0 in the low half.
template <int VR>
__device__ __forceinline__ void load_vr(const bf16_bits* p, float (&x)[VR]) {
  if constexpr (VR % 4 == 0) {
#pragma unroll
    for (int h = 0; h < VR; h += 4) {
      const uint2 t = *reinterpret_cast<const uint2*>(p + h);
      x[h] = __uint_as_float(t.x << 16);
      x[h + 1] = __uint_as_float(t.x & 0xffff0000u);
      x[h + 2] = __uint_as_float(t.y << 16);
      x[h + 3] = __uint_as_float(t.y & 0xffff0000u);
    }
  } else {
#pragma unroll
    for (int q = 0; q < VR; ++q) x[q] = bf16_to_f32(p[q]);
  }
}

template <int VR>
__device__ __forceinline__ void store_vr(bf16_bits* p, const float (&x)[VR]) {
  if constexpr (VR % 4 == 0) {
#pragma unroll
    for (int h = 0; h < VR; h += 4) {
      uint2 t;
      t.x = (uint32_t)f32_to_bf16(x[h]) | ((uint32_t)f32_to_bf16(x[h + 1]) << 16);
      t.y = (uint32_t)f32_to_bf16(x[h + 2]) | ((uint32_t)f32_to_bf16(x[h + 3]) << 16);
      *reinterpret_cast<uint2*>(p + h) = t;
    }
  } else {
#pragma unroll
    for (int q = 0; q < VR; ++q) p[q] = f32_to_bf16(x[q]);
  }
}

// Back projection with VR detector rows per lane (R % 4 == 0): a pixel's
// cylinder is LPX lanes x VR rows, a wavefront holds 64/LPX pixel slots
// (consecutive list entries, so neighbouring pixels of an image row share
// tap rows in the same load instruction), and each lane accumulates ACC
// pixels.  Every tap is one float4 (or 4 x bfloat16) load per lane with a
// per-lane address: no wave-uniform window, no scalar address arithmetic.
// Geometry for (view batch x pixels) in LDS as in par_back_kernel; the
// same view splits and the same fixed-order split reduction.
template <int LPX, int VR, int ACC, int NT, int CP, bool WEIGHTED, typename TS>
__global__ __launch_bounds__(kBlock) void par_back_vr_kernel(
    const TS* __restrict__ sino, const TS* __restrict__ wts,
    const int64_t* __restrict__ pix, int n_pix, const ParView* __restrict__ views,
    mbir_par_geom g, int views_per_split, float* __restrict__ out) {
  constexpr int SL = MBIR_WAVE / LPX;  // pixel slots per wavefront
  constexpr int PIX_WAVE = SL * ACC;
  constexpr int PIX_WG = (kBlock / MBIR_WAVE) * PIX_WAVE;
  constexpr int VB = PIX_WG >= 2048 ? 1 : 2048 / PIX_WG;  // views per LDS batch
  constexpr int ROWS = LPX * VR;
  __shared__ int s_n0[VB * PIX_WG];
  __shared__ float s_a[NT * VB * PIX_WG];

  const int lane = threadIdx.x & (MBIR_WAVE - 1), wave = threadIdx.x / MBIR_WAVE;
  const int slot = lane / LPX, rl = lane % LPX;
  const int R = g.num_slices, C = g.num_channels;
  int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
  if (gridDim.y > 1) {  // split view ranges: one XCD runs one split's pixel groups
    const int nx = gridDim.x, ny = gridDim.y;
    const int l = xcd_remap(blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z), nx * ny * gridDim.z);
    bx = l % nx;
    by = (l / nx) % ny;
    bz = l / (nx * ny);
  }
  const int rr = bz * ROWS + rl * VR;
  const bool r_ok = rr < R;  // R % 4 == 0: rows rr .. rr+VR-1 all exist
  const int p_base = bx * PIX_WG;
  const int v_begin = by * views_per_split;
  const int v_end = min(v_begin + views_per_split, g.num_views);

  float acc[ACC][VR];
#pragma unroll
  for (int k = 0; k < ACC; ++k)
#pragma unroll
    for (int q = 0; q < VR; ++q) acc[k][q] = 0.0f;

  for (int vb0 = v_begin; vb0 < v_end; vb0 += VB) {
    const int nvb = min(VB, v_end - vb0);
    __syncthreads();
    for (int q = threadIdx.x; q < VB * PIX_WG; q += kBlock) {
      const int j = q % PIX_WG, b = q / PIX_WG;
      const int pg = p_base + j;
      int base = 0;
      float a[NT];
      if (b < nvb && pg < n_pix) {
        par_taps<NT>(pix[pg], views[vb0 + b], g, base, a);
      } else {
#pragma unroll
        for (int k = 0; k < NT; ++k) a[k] = 0.0f;
      }
      s_n0[q] = base;
#pragma unroll
      for (int k = 0; k < NT; ++k) s_a[k * VB * PIX_WG + q] = CP == 2 ? a[k] * a[k] : a[k];
    }
    __syncthreads();
    if (r_ok) {
      for (int b = 0; b < nvb; ++b) {
        const size_t vofs = (size_t)(vb0 + b) * C * R + rr;
#pragma unroll
        for (int k = 0; k < ACC; ++k) {
          const int q = b * PIX_WG + wave * PIX_WAVE + k * SL + slot;
          const int base = s_n0[q];
          float x[NT][VR];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            // Off-detector taps carry weight 0; the clamp keeps the load in bounds.
            const size_t o = vofs + (size_t)min(max(base + t, 0), C - 1) * R;
            load_vr<VR>(sino + o, x[t]);
            if (WEIGHTED) {
              float w[VR];
              load_vr<VR>(wts + o, w);
#pragma unroll
              for (int r = 0; r < VR; ++r) x[t][r] = x[t][r] * w[r];
            }
          }
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const float a = s_a[t * VB * PIX_WG + q];
#pragma unroll
            for (int r = 0; r < VR; ++r) acc[k][r] = fmaf(a, x[t][r], acc[k][r]);
          }
        }
      }
    }
  }
  if (r_ok) {
#pragma unroll
    for (int k = 0; k < ACC; ++k) {
      const int pg = p_base + wave * PIX_WAVE + k * SL + slot;
      if (pg < n_pix) store_vr<VR>(out + ((size_t)by * n_pix + pg) * R + rr, acc[k]);
    }
  }
}

enum FwdMode { FWD_WRITE_VCR = 0, FWD_WRITE_VRC = 1, FWD_LINESEARCH = 2, FWD_UPDATE = 3 };

// Channel-major epilogue of one lane: sinogram entries o .. o+VR-1 hold
// (v, n, rr .. rr+VR-1).  Line search: FP64 sums of (w*e)*t and (t*t)*w;
// update: e -= alpha*t.
template <int VR, int MODE, bool WEIGHTED, typename TS>
__device__ __forceinline__ void fwd_epilogue_vcr(const float (&t)[VR], size_t o,
                                                 TS* __restrict__ sino,
                                                 const TS* __restrict__ wts, float al,
                                                 float* __restrict__ adelta,
                                                 double& lin, double& quad) {
  if (MODE == FWD_WRITE_VCR) {
    if constexpr (sizeof(TS) == sizeof(float)) store_vr<VR>(sino + o, t);
  } else if (MODE == FWD_LINESEARCH) {
    if (adelta) store_vr<VR>(adelta + o, t);  // A(delta) kept for the error update
    // No error sinogram: the linear term <W e, A delta> comes from the back
    // projection instead, and only the quadratic term is formed here.
    float e[VR];
    if (sino) {
      load_vr<VR>(sino + o, e);
    } else {
#pragma unroll
      for (int q = 0; q < VR; ++q) e[q] = 0.0f;
    }
    if (WEIGHTED) {
      float w[VR];
      load_vr<VR>(wts + o, w);
#pragma unroll
      for (int q = 0; q < VR; ++q) {
        if (sino) lin += (double)((w[q] * e[q]) * t[q]);
        quad += (double)((t[q] * t[q]) * w[q]);
      }
    } else {
#pragma unroll
      for (int q = 0; q < VR; ++q) {
        if (sino) lin += (double)(e[q] * t[q]);
        quad += (double)(t[q] * t[q]);
      }
    }
  } else if (MODE == FWD_UPDATE) {
    float e[VR];
    load_vr<VR>(sino + o, e);
#pragma unroll
    for (int q = 0; q < VR; ++q) e[q] = e[q] - al * t[q];
    store_vr<VR>(sino + o, e);
  }
}

// Line-search partials: one (lin, quad) pair per workgroup at its linear
// block index.
__device__ __forceinline__ void fwd_write_partials(double lin, double quad,
                                                   double* __restrict__ partials) {
  block_sum2_f64(lin, quad);
  const int b = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  if (threadIdx.x == 0) {
    partials[2 * b] = lin;
    partials[2 * b + 1] = quad;
  }
}

// ---------------------------------------------------------------------------
// Ray-driven forward for row-structured pixel lists.  runs[3*i .. 3*i+2] =
// (list position of the run's first pixel, first column, run length) for
// image row i; a length of 0 means the row holds no listed pixel.
//
// Workgroup = (CHW channels, one view, one row chunk); a unit owns CPS
// adjacent channels n0 .. n0+CPS-1 and walks the image rows (or columns,
// whichever the view's projection is steeper in).  The candidate pixels of
// each image row are the columns whose footprint can reach the unit's
// channels, from the inverse map with a 0.01-pixel margin; each candidate's
// n_p is evaluated once with the exact chain and its tap weights for the
// CPS channels follow the same rule as the back projector, so every nonzero
// (pixel, channel) weight is found and weights agree bit for bit.
constexpr float kRayMargin = 0.01f;

// NB candidates (i[u], j[u]) at list positions pos[u], those with ok[u]
// live, in one batch: geometry first, then all cylinder loads, then the FMAs,
// so the loads of a batch are in flight together.
template <int NB, int VR, int CPS, bool TWO>
__device__ __forceinline__ void ray_taps(const int (&ci)[NB], const int (&cj)[NB],
                                         const int (&pos)[NB], const bool (&ok)[NB], int n0,
                                         int ntaps,
                                          bool row_ok, int rr, const ParView& vw,
                                          const mbir_par_geom& g,
                                          const float* __restrict__ values,
                                          float (&acc)[CPS][VR]) {
  // The trapezoid weight is zero outside the pixel's taps (its support
  // |n - n_p| < (W+1)/2 lies inside round(n_p) +- psf, and inside floor(n_p),
  // floor(n_p)+1 when W <= 1), so the weight alone selects the taps.
  float a[NB][CPS];
  bool use[NB];
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const float n_p = par_np(ci[u], cj[u], vw, g);
    bool any = false;
#pragma unroll
    for (int k = 0; k < CPS; ++k) {
      a[u][k] = par_weight(n_p, n0 + k, vw);
      any = any || a[u][k] != 0.0f;
    }
    use[u] = any && row_ok && ok[u];
  }
  float x[NB][VR];
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    if (use[u]) {
      load_vr<VR>(values + (size_t)pos[u] * g.num_slices + rr, x[u]);
    } else {
#pragma unroll
      for (int q = 0; q < VR; ++q) x[u][q] = 0.0f;
    }
  }
#pragma unroll
  for (int u = 0; u < NB; ++u)
#pragma unroll
    for (int k = 0; k < CPS; ++k)
#pragma unroll
      for (int q = 0; q < VR; ++q) acc[k][q] = fmaf(a[u][k], x[u][q], acc[k][q]);
}

// Four-lane units: each lane evaluates one candidate of a batch of four
// (n_p and its CPS weights, the list position, live or not), and the unit's
// lanes exchange them with quad-permute DPP moves, so the geometry is
// computed once per candidate instead of once per lane.  The loads of the
// batch go out in PH phases of 4/PH candidates (register budget).
template <int U>
__device__ __forceinline__ int quad_bcast(int x) {
  return __builtin_amdgcn_mov_dpp(x, U | (U << 2) | (U << 4) | (U << 6), 0xF, 0xF, false);
}
template <int U>
__device__ __forceinline__ float quad_bcast(float x) {
  return __int_as_float(quad_bcast<U>(__float_as_int(x)));
}

template <int U, int CPS>
__device__ __forceinline__ void quad_take(const float (&a)[CPS], int pos, int use,
                                          float (&A)[4][CPS], int (&POS)[4], int (&USE)[4]) {
#pragma unroll
  for (int k = 0; k < CPS; ++k) A[U][k] = quad_bcast<U>(a[k]);
  POS[U] = quad_bcast<U>(pos);
  USE[U] = quad_bcast<U>(use);
}

template <int VR, int CPS, int PH>
__device__ __forceinline__ void quad_ray_taps(int ci, int cj, int pos, bool ok, int n0,
                                              bool row_ok, int rr, const ParView& vw,
                                              const mbir_par_geom& g,
                                              const float* __restrict__ values,
                                              float (&acc)[CPS][VR]) {
  const float n_p = par_np(ci, cj, vw, g);
  float a[CPS];
  bool any = false;
#pragma unroll
  for (int k = 0; k < CPS; ++k) {
    a[k] = par_weight(n_p, n0 + k, vw);
    any = any || a[k] != 0.0f;
  }
  float A[4][CPS];
  int POS[4], USE[4];
  const int use = (ok && any) ? 1 : 0;
  quad_take<0, CPS>(a, pos, use, A, POS, USE);
  quad_take<1, CPS>(a, pos, use, A, POS, USE);
  quad_take<2, CPS>(a, pos, use, A, POS, USE);
  quad_take<3, CPS>(a, pos, use, A, POS, USE);
  constexpr int PER = 4 / PH;
#pragma unroll
  for (int h = 0; h < PH; ++h) {
    float x[PER][VR];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      if (USE[h * PER + u] && row_ok) {
        load_vr<VR>(values + (size_t)POS[h * PER + u] * g.num_slices + rr, x[u]);
      } else {
#pragma unroll
        for (int q = 0; q < VR; ++q) x[u][q] = 0.0f;
      }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u)
#pragma unroll
      for (int k = 0; k < CPS; ++k)
#pragma unroll
        for (int q = 0; q < VR; ++q) acc[k][q] = fmaf(A[h * PER + u][k], x[u][q], acc[k][q]);
  }
}

template <int LPC, int VR, int CPS, int NB, bool TWO, int MODE, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void par_fwd_ray_kernel(
    const float* __restrict__ values, const int* __restrict__ runs,
    const ParView* __restrict__ views, mbir_par_geom g, float* __restrict__ sino,
    const float* __restrict__ wts, const float* __restrict__ alpha,
    double* __restrict__ partials, float* __restrict__ adelta) {
  using L = FwdLanes<LPC, VR>;
  constexpr int CHW = L::UNITS * CPS;
  constexpr int TS = CHW + 1;
  // Four-lane units with NB a multiple of 4 share candidate geometry across
  // the unit (NB / 4 load phases per batch of four).
  constexpr bool QUAD = LPC == 4 && NB % 4 == 0;
  constexpr int QPH = QUAD ? NB / 4 : 1;
  __shared__ float s_tile[MODE == FWD_WRITE_VRC ? L::ROWS * TS : 1];

  const int lane = threadIdx.x & (MBIR_WAVE - 1), wave = threadIdx.x / MBIR_WAVE;
  const int unit = wave * L::SPW + lane / LPC, rl = lane % LPC;
  // Logical order: view fastest, so an XCD runs one channel block over
  // consecutive views (overlapping pixel bands share its L2).
  int c_base, v, r0;
  {
    const int nv = gridDim.y, ncb = gridDim.x;
    const int l = xcd_remap(blockIdx.x + ncb * (blockIdx.y + nv * blockIdx.z),
                            ncb * nv * gridDim.z);
    v = l % nv;
    c_base = ((l / nv) % ncb) * CHW;
    r0 = (l / (nv * ncb)) * L::ROWS;
  }
  const int R = g.num_slices, C = g.num_channels, Nr = g.num_rows, Nc = g.num_cols;
  const int n0 = c_base + unit * CPS;
  const int rr = r0 + rl * VR;
  const bool row_ok = rr < R;
  const int ntaps = TWO ? 2 : 2 * g.psf_radius + 1;
  const ParView vw = views[v];

  float acc[CPS][VR];
#pragma unroll
  for (int k = 0; k < CPS; ++k)
#pragma unroll
    for (int q = 0; q < VR; ++q) acc[k][q] = 0.0f;

  const float ddc = g.delta_det_channel;
  const float xa = ((float)n0 - g.det_center) * ddc - g.det_channel_offset;
  const float xb = ((float)(n0 + CPS - 1) - g.det_center) * ddc - g.det_channel_offset;
  const float cd = vw.cosv * g.delta_voxel, sd = vw.sinv * g.delta_voxel_row;
  // The candidate ranges use reciprocals: their rounding (~1e-7 relative) is
  // far inside the 0.01-pixel margin; each candidate's n_p is exact.
  if (fabsf(cd) >= fabsf(sd)) {
    const float hj = vw.halfw * ddc / fabsf(cd) + kRayMargin;
    const float icd = 1.0f / cd;
    for (int i = 0; i < Nr; ++i) {
      const int start = runs[3 * i], jmin = runs[3 * i + 1], cnt = runs[3 * i + 2];
      if (cnt <= 0) continue;
      const float sy = vw.sinv * (g.delta_voxel_row * ((float)i - g.row_center));
      const float ja = (xa + sy) * icd + g.col_center, jb = (xb + sy) * icd + g.col_center;
      const int jlo = max((int)floorf(fminf(ja, jb) - hj), jmin);
      const int jhi = min((int)ceilf(fmaxf(ja, jb) + hj), jmin + cnt - 1);
      if constexpr (QUAD) {
        for (int j = jlo; j <= jhi; j += 4) {
          const int jj = min(j + rl, jhi);
          quad_ray_taps<VR, CPS, QPH>(i, jj, start + jj - jmin, j + rl <= jhi, n0, row_ok, rr,
                                      vw, g, values, acc);
        }
      } else {
        for (int j = jlo; j <= jhi; j += NB) {
          int ci[NB], cj[NB], pos[NB];
          bool ok[NB];
#pragma unroll
          for (int u = 0; u < NB; ++u) {
            const int jj = min(j + u, jhi);
            ci[u] = i;
            cj[u] = jj;
            pos[u] = start + jj - jmin;
            ok[u] = j + u <= jhi;
          }
          ray_taps<NB, VR, CPS, TWO>(ci, cj, pos, ok, n0, ntaps, row_ok, rr, vw, g, values, acc);
        }
      }
    }
  } else {
    const float hi = vw.halfw * ddc / fabsf(sd) + kRayMargin;
    const float isd = 1.0f / sd;
    for (int j = 0; j < Nc; ++j) {
      const float cx = vw.cosv * (g.delta_voxel * ((float)j - g.col_center));
      const float ia = (cx - xa) * isd + g.row_center, ib = (cx - xb) * isd + g.row_center;
      const int ilo = max((int)floorf(fminf(ia, ib) - hi), 0);
      const int ihi = min((int)ceilf(fmaxf(ia, ib) + hi), Nr - 1);
      if constexpr (QUAD) {
        for (int i = ilo; i <= ihi; i += 4) {
          const int ii = min(i + rl, ihi);
          const int start = runs[3 * ii], jmin = runs[3 * ii + 1], cnt = runs[3 * ii + 2];
          quad_ray_taps<VR, CPS, QPH>(ii, j, start + j - jmin,
                                      i + rl <= ihi && j >= jmin && j < jmin + cnt, n0, row_ok,
                                      rr, vw, g, values, acc);
        }
      } else {
        for (int i = ilo; i <= ihi; i += NB) {
          int ci[NB], cj[NB], pos[NB];
          bool ok[NB];
#pragma unroll
          for (int u = 0; u < NB; ++u) {
            const int ii = min(i + u, ihi);
            const int start = runs[3 * ii], jmin = runs[3 * ii + 1], cnt = runs[3 * ii + 2];
            ci[u] = ii;
            cj[u] = j;
            pos[u] = start + j - jmin;
            ok[u] = i + u <= ihi && j >= jmin && j < jmin + cnt;
          }
          ray_taps<NB, VR, CPS, TWO>(ci, cj, pos, ok, n0, ntaps, row_ok, rr, vw, g, values, acc);
        }
      }
    }
  }

  if (MODE == FWD_WRITE_VRC) {
    // Stage the (ROWS x CHW) tile so each detector row is stored contiguously.
    if (row_ok) {
#pragma unroll
      for (int k = 0; k < CPS; ++k)
#pragma unroll
        for (int q = 0; q < VR; ++q) s_tile[(rl * VR + q) * TS + unit * CPS + k] = acc[k][q];
    }
    __syncthreads();
    const int nrow = min(L::ROWS, R - r0), nch = min(CHW, C - c_base);
    for (int e = threadIdx.x; e < L::ROWS * CHW; e += kBlock) {
      const int row = e / CHW, c = e - row * CHW;
      if (row < nrow && c < nch)
        sino[((size_t)v * R + r0 + row) * C + c_base + c] = s_tile[row * TS + c];
    }
  } else {
    double lin = 0.0, quad = 0.0;
    const float al = MODE == FWD_UPDATE ? alpha[0] : 0.0f;
    if (row_ok) {
#pragma unroll
      for (int k = 0; k < CPS; ++k) {
        const int n = n0 + k;
        if (n < C)
          fwd_epilogue_vcr<VR, MODE, WEIGHTED>(acc[k], ((size_t)v * C + n) * R + rr, sino, wts,
                                               al, adelta, lin, quad);
      }
    }
    if (MODE == FWD_LINESEARCH) fwd_write_partials(lin, quad, partials);
  }
}

// ---------------------------------------------------------------------------
// Binned forward for arbitrary pixel lists (the VCD subsets).  Workgroup =
// (view, CHW channels, row chunk); units own CPS adjacent channels as in the
// ray form.  Each chunk of the list is counting-sorted in LDS by the first
// tap channel of its pixels -- one entry per pixel: list position, bin and
// the two tap weights (two-tap rule) or n_p (2*psf+1 taps) -- bins ordered by
// list position.  The pixels whose taps reach a unit's channels n0 ..
// n0+CPS-1 are then one contiguous segment of the sorted chunk (first taps
// n0-ntaps+1 .. n0+CPS-1), which the unit walks NB candidates at a time:
// tap weights for its CPS channels, the candidates' cylinder loads in flight
// together, then the FMAs.  A pixel's cylinder is loaded by the one or two
// units its taps reach.  The summation order of every sinogram entry is
// fixed by the list: deterministic.
constexpr int kSegChunk = 2048;   // pixels per chunk (positions fit 16 bits)
constexpr int kSegShortBin = 32;  // longest bin ranked by counting

// Exclusive block scan of one int per thread (256 threads).
__device__ __forceinline__ int block_exclusive_scan(int v, int* s_tmp) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_tmp[wid] = x;
  __syncthreads();
  int prefix = 0;
  for (int w = 0; w < wid; ++w) prefix += s_tmp[w];
  return prefix + x - v;
}

template <int LPC, int VR, int CPS, int NB, bool TWO, int MODE, bool WEIGHTED, typename TS>
__global__ __launch_bounds__(kBlock) void par_fwd_bin_kernel(
    const float* __restrict__ values, const int64_t* __restrict__ pix, int n_pix,
    const ParView* __restrict__ views, mbir_par_geom g, TS* __restrict__ sino,
    const TS* __restrict__ wts, const float* __restrict__ alpha,
    double* __restrict__ partials, float* __restrict__ adelta) {
  using L = FwdLanes<LPC, VR>;
  constexpr int CHW = L::UNITS * CPS;
  constexpr int NBIN_MAX = CHW + (TWO ? 1 : kMaxTaps - 1);
  constexpr int PPT = kSegChunk / kBlock;
  constexpr int BPT = (NBIN_MAX + kBlock - 1) / kBlock;  // bins per thread in the scan
  __shared__ int s_cnt[BPT * kBlock];
  __shared__ int s_off[BPT * kBlock + 1];
  __shared__ int s_tmp[4];
  __shared__ int s_pk[kSegChunk];  // list position | bin << 16
  __shared__ int s_wa[kSegChunk];  // float bits: first tap weight (two-tap) or n_p
  __shared__ int s_wb[TWO ? kSegChunk : 1];  // float bits: second tap weight
  // Unsorted list positions of the short-bin placement, in the weight array
  // that the sorted entries fill only after every rank is known.
  int* const s_pos2 = TWO ? s_wb : s_wa;
  __shared__ int s_wcnt[kBlock / MBIR_WAVE][NBIN_MAX];

  const int tid = threadIdx.x;
  const int lane = tid & (MBIR_WAVE - 1), wave = tid / MBIR_WAVE;
  const int unit = wave * L::SPW + lane / LPC, rl = lane % LPC;
  const int v = blockIdx.x, c0 = blockIdx.y * CHW, r0 = blockIdx.z * L::ROWS;
  const int R = g.num_slices, C = g.num_channels, Nc = g.num_cols;
  const int rr = r0 + rl * VR;
  const bool row_ok = rr < R;
  const int ntaps = TWO ? 2 : 2 * g.psf_radius + 1;
  const int nbin = CHW + ntaps - 1;
  const int lo = c0 - (ntaps - 1);  // first-tap channel of bin 0
  const int n0 = c0 + unit * CPS;
  const int tb0 = unit * CPS + ntaps - 1;  // bin whose first tap is n0
  const ParView vw = views[v];
  for (int i = tid; i < (kBlock / MBIR_WAVE) * NBIN_MAX; i += kBlock) (&s_wcnt[0][0])[i] = 0;

  float acc[CPS][VR];
#pragma unroll
  for (int k = 0; k < CPS; ++k)
#pragma unroll
    for (int q = 0; q < VR; ++q) acc[k][q] = 0.0f;

  for (int pc = 0; pc < n_pix; pc += kSegChunk) {
    int bin[PPT];
    float npv[PPT];
#pragma unroll
    for (int i = 0; i < BPT; ++i) s_cnt[tid * BPT + i] = 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      const int j = pc + tid + q * kBlock;
      int b = -1;
      float n_p = 0.0f;
      if (j < n_pix) {
        const int pj = (int)pix[j];  // pixel indices < rows*cols < 2^31
        const int row = pj / Nc, col = pj - row * Nc;
        n_p = par_np(row, col, vw, g);
        b = par_base<TWO ? 2 : 0>(n_p, g) - lo;
        if (b < 0 || b >= nbin) b = -1;
      }
      bin[q] = b;
      npv[q] = n_p;
      if (b >= 0) atomicAdd(&s_cnt[b], 1);
    }
    __syncthreads();
    // Bins tid*BPT .. tid*BPT+BPT-1 (counts zero beyond nbin).
    int cnt[BPT], mine = 0, longest = 0;
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
      cnt[i] = s_cnt[tid * BPT + i];
      mine += cnt[i];
      longest = max(longest, cnt[i]);
    }
    int off = block_exclusive_scan(mine, s_tmp);
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
      s_off[tid * BPT + i] = off;
      s_cnt[tid * BPT + i] = off;  // placement cursor of the bin
      off += cnt[i];
    }
    if (tid == kBlock - 1) s_off[BPT * kBlock] = off;
    // Entry of slot q at its sorted place sl.
    auto put = [&](int q, int sl) {
      const int me = tid + q * kBlock;
      s_pk[sl] = me | (bin[q] << 16);
      if constexpr (TWO) {
        const int base = lo + bin[q];
        s_wa[sl] = __float_as_int(par_weight(npv[q], base, vw));
        s_wb[sl] = __float_as_int(par_weight(npv[q], base + 1, vw));
      } else {
        s_wa[sl] = __float_as_int(npv[q]);
      }
    };
    if (!__syncthreads_or(longest > kSegShortBin)) {
      // Short bins: place with LDS atomics, then move each entry to its rank
      // by list position within the bin (positions in a chunk are distinct).
#pragma unroll
      for (int q = 0; q < PPT; ++q)
        if (bin[q] >= 0) s_pos2[atomicAdd(&s_cnt[bin[q]], 1)] = tid + q * kBlock;
      __syncthreads();
      int slot[PPT];
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        slot[q] = -1;
        if (bin[q] >= 0) {
          const int e0 = s_off[bin[q]], e1 = s_off[bin[q] + 1];
          const int me = tid + q * kBlock;
          int rank = 0;
          for (int e = e0; e < e1; ++e) rank += s_pos2[e] < me;
          slot[q] = e0 + rank;
        }
      }
      __syncthreads();  // every rank read before the weights overwrite s_pos2
#pragma unroll
      for (int q = 0; q < PPT; ++q)
        if (slot[q] >= 0) put(q, slot[q]);
    } else {
      // Long bins (a dense list: an image row onto one channel): stable
      // placement round by round (pixel slot q) and within a round by thread
      // id -- the bin cursor, plus the same-bin entries of lower waves, plus
      // the rank among same-bin lanes below in the wave.
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        const int b = bin[q];
        int rank, gsize;
        bool leader;
        wave_bin_rank(b, rank, gsize, leader);
        if (leader) s_wcnt[wave][b] = gsize;
        __syncthreads();
        if (b >= 0) {
          int sl = s_cnt[b] + rank;
          for (int w = 0; w < wave; ++w) sl += s_wcnt[w][b];
          put(q, sl);
        }
        __syncthreads();
        for (int bb = tid; bb < nbin; bb += kBlock) {
          int add = 0;
#pragma unroll
          for (int w = 0; w < kBlock / MBIR_WAVE; ++w) {
            add += s_wcnt[w][bb];
            s_wcnt[w][bb] = 0;
          }
          s_cnt[bb] += add;
        }
        __syncthreads();
      }
    }
    __syncthreads();
    if (row_ok) {
      const float* xb = values + (size_t)pc * R + rr;
      const int b0 = unit * CPS;  // first taps n0-ntaps+1 .. n0+CPS-1
      const int s0 = s_off[b0], s1 = s_off[min(b0 + CPS + ntaps - 1, nbin)];
      for (int e = s0; e < s1; e += NB) {
        float a[NB][CPS];
        int pos[NB];
        bool use[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          const bool ok = e + u < s1;
          const int ee = ok ? e + u : s0;
          const int pk = s_pk[ee];
          pos[u] = pk & 0xffff;
          bool any = false;
          if constexpr (TWO) {
            // Taps n0+t and n0+t+1 carry the stored weights.
            const int t = (pk >> 16) - tb0;
            const float w0 = __int_as_float(s_wa[ee]), w1 = __int_as_float(s_wb[ee]);
#pragma unroll
            for (int k = 0; k < CPS; ++k) {
              a[u][k] = t == k ? w0 : (t == k - 1 ? w1 : 0.0f);
              any = any || a[u][k] != 0.0f;
            }
          } else {
            const float n_p = __int_as_float(s_wa[ee]);
            const int t0 = (pk >> 16) - tb0;  // first tap - n0
#pragma unroll
            for (int k = 0; k < CPS; ++k) {
              const int t = k - t0;
              a[u][k] = (t >= 0 && t < ntaps) ? par_weight(n_p, n0 + k, vw) : 0.0f;
              any = any || a[u][k] != 0.0f;
            }
          }
          use[u] = ok && any;
        }
        float x[NB][VR];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          if (use[u]) {
            load_vr<VR>(xb + (size_t)pos[u] * R, x[u]);
          } else {
#pragma unroll
            for (int q = 0; q < VR; ++q) x[u][q] = 0.0f;
          }
        }
#pragma unroll
        for (int u = 0; u < NB; ++u)
#pragma unroll
          for (int k = 0; k < CPS; ++k)
#pragma unroll
            for (int q = 0; q < VR; ++q) acc[k][q] = fmaf(a[u][k], x[u][q], acc[k][q]);
      }
    }
    __syncthreads();
  }

  double lin = 0.0, quad = 0.0;
  const float al = MODE == FWD_UPDATE ? alpha[0] : 0.0f;
  if (row_ok) {
#pragma unroll
    for (int k = 0; k < CPS; ++k) {
      const int n = n0 + k;
      if (n >= C) continue;
      if (MODE == FWD_WRITE_VRC) {
        if constexpr (sizeof(TS) == sizeof(float)) {
#pragma unroll
          for (int q = 0; q < VR; ++q)
            if (rr + q < R) sino[((size_t)v * R + rr + q) * C + n] = acc[k][q];
        }
      } else {
        fwd_epilogue_vcr<VR, MODE, WEIGHTED>(acc[k], ((size_t)v * C + n) * R + rr, sino, wts,
                                             al, adelta, lin, quad);
      }
    }
  }
  if (MODE == FWD_LINESEARCH) fwd_write_partials(lin, quad, partials);
}

// ---------------------------------------------------------------------------
// Host-side launch helpers.
static inline int next_pow2(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

static int check_geom(const mbir_par_geom* g) {
  if (!g) return MBIR_ERR_ARG;
  if (g->num_rows <= 0 || g->num_cols <= 0 || g->num_slices <= 0 ||
      g->num_channels <= 0 || g->num_views <= 0)
    return MBIR_ERR_SHAPE;
  if (g->psf_radius < 0 || 2 * g->psf_radius + 1 > kMaxTaps) return MBIR_ERR_GEOMETRY;
  if (!(g->delta_det_channel > 0.0f) || !(g->delta_voxel > 0.0f) || !(g->delta_voxel_row > 0.0f))
    return MBIR_ERR_GEOMETRY;
  return MBIR_OK;
}

static inline int back_rw(int R) { return R >= 64 ? 64 : next_pow2(R); }

static inline int back_pix_per_wg(int RW) {
  return (kBlock / MBIR_WAVE) * (MBIR_WAVE / RW) * (RW >= 4 ? 8 : 4);
}

// Sinogram storage: float32, or (cp == 1 only) bfloat16 bits.
template <int RW, int NT>
static int launch_back_rw(const void* sino, const void* w, const int64_t* pix, int P,
                          const ParView* views, const mbir_par_geom& g, int cp, int ns,
                          float* out, hipStream_t s, bool bf16) {
  const int pwg = back_pix_per_wg(RW);
  const int vps = (g.num_views + ns - 1) / ns;
  dim3 grid((P + pwg - 1) / pwg, ns, (g.num_slices + RW - 1) / RW);
  const float* sf = static_cast<const float*>(sino);
  const float* wf = static_cast<const float*>(w);
  const bf16_bits* sh = static_cast<const bf16_bits*>(sino);
  const bf16_bits* wh = static_cast<const bf16_bits*>(w);
  if (bf16) {
    if (w) par_back_kernel<RW, NT, 1, true, bf16_bits><<<grid, kBlock, 0, s>>>(sh, wh, pix, P, views, g, vps, out);
    else par_back_kernel<RW, NT, 1, false, bf16_bits><<<grid, kBlock, 0, s>>>(sh, wh, pix, P, views, g, vps, out);
  } else if (cp == 1) {
    if (w) par_back_kernel<RW, NT, 1, true, float><<<grid, kBlock, 0, s>>>(sf, wf, pix, P, views, g, vps, out);
    else par_back_kernel<RW, NT, 1, false, float><<<grid, kBlock, 0, s>>>(sf, wf, pix, P, views, g, vps, out);
  } else {
    if (w) par_back_kernel<RW, NT, 2, true, float><<<grid, kBlock, 0, s>>>(sf, wf, pix, P, views, g, vps, out);
    else par_back_kernel<RW, NT, 2, false, float><<<grid, kBlock, 0, s>>>(sf, wf, pix, P, views, g, vps, out);
  }
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

// Row-vector back projection (R % 4 == 0, 16-byte aligned rows): lanes per
// pixel and pixels per lane by row count.  Measured at 720 views x 64 rows:
// 16 lanes x 4 pixels per lane (64 pixels per workgroup) 2.6 ms for the full
// 512x512 grid against 3.6 ms for par_back_kernel, and 59-1311 us against
// 63-1419 us for 1,602-51,253-pixel weighted subsets.
struct BackVrCfg {
  int lpx, acc;
};
static inline BackVrCfg back_vr_cfg(int R) {
  if (R >= 64) return {16, 4};
  if (R >= 32) return {8, 4};
  if (R >= 16) return {4, 2};
  if (R >= 8) return {2, 1};
  return {1, 1};
}
static inline bool back_vr_ok(const mbir_par_geom& g) { return g.num_slices % 4 == 0; }
// Pixels per workgroup and rows per chunk of the form mbir_par_back takes.
static inline void back_shape(const mbir_par_geom& g, int& pwg, int& rows) {
  if (back_vr_ok(g)) {
    const BackVrCfg c = back_vr_cfg(g.num_slices);
    pwg = (kBlock / MBIR_WAVE) * (MBIR_WAVE / c.lpx) * c.acc;
    rows = c.lpx * 4;
  } else {
    rows = back_rw(g.num_slices);
    pwg = back_pix_per_wg(rows);
  }
}

template <int LPX, int ACC, int NT>
static int launch_back_vr(const void* sino, const void* w, const int64_t* pix, int P,
                          const ParView* views, const mbir_par_geom& g, int cp, int ns,
                          float* out, hipStream_t s, bool bf16) {
  constexpr int PWG = (kBlock / MBIR_WAVE) * (MBIR_WAVE / LPX) * ACC;
  constexpr int ROWS = LPX * 4;
  const int vps = (g.num_views + ns - 1) / ns;
  dim3 grid((P + PWG - 1) / PWG, ns, (g.num_slices + ROWS - 1) / ROWS);
  const float* sf = static_cast<const float*>(sino);
  const float* wf = static_cast<const float*>(w);
  const bf16_bits* sh = static_cast<const bf16_bits*>(sino);
  const bf16_bits* wh = static_cast<const bf16_bits*>(w);
#define MBIR_BACK_VR(CP, W, T, S, WT) \
  par_back_vr_kernel<LPX, 4, ACC, NT, CP, W, T><<<grid, kBlock, 0, s>>>(S, WT, pix, P, views, g, vps, out)
  if (bf16) {
    if (w) MBIR_BACK_VR(1, true, bf16_bits, sh, wh);
    else MBIR_BACK_VR(1, false, bf16_bits, sh, wh);
  } else if (cp == 1) {
    if (w) MBIR_BACK_VR(1, true, float, sf, wf);
    else MBIR_BACK_VR(1, false, float, sf, wf);
  } else {
    if (w) MBIR_BACK_VR(2, true, float, sf, wf);
    else MBIR_BACK_VR(2, false, float, sf, wf);
  }
#undef MBIR_BACK_VR
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

template <int NT>
static int launch_back_nt(const void* sino, const void* w, const int64_t* pix, int P,
                          const ParView* views, const mbir_par_geom& g, int cp, int ns,
                          float* out, hipStream_t s, bool bf16) {
  const uintptr_t align = bf16 ? 7 : 15;
  if (back_vr_ok(g) && (reinterpret_cast<uintptr_t>(sino) & align) == 0 &&
      (!w || (reinterpret_cast<uintptr_t>(w) & align) == 0) &&
      (reinterpret_cast<uintptr_t>(out) & 15) == 0) {
    switch (back_vr_cfg(g.num_slices).lpx) {
      case 16: return launch_back_vr<16, 4, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
      case 8: return launch_back_vr<8, 4, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
      case 4: return launch_back_vr<4, 2, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
      case 2: return launch_back_vr<2, 1, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
      default: return launch_back_vr<1, 1, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
    }
  }
  switch (back_rw(g.num_slices)) {
    case 64: return launch_back_rw<64, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
    case 32: return launch_back_rw<32, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
    case 16: return launch_back_rw<16, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
    case 8: return launch_back_rw<8, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
    case 4: return launch_back_rw<4, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
    case 2: return launch_back_rw<2, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
    default: return launch_back_rw<1, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
  }
}
// Forward dispatch: the ray form when `runs` is given, the binned form
// otherwise.  Lane configuration: VR = 4 rows per lane (float4) when R % 4 == 0
// and every row array is 16-byte aligned, else VR = 1.
struct FwdArgs {
  int mode;
  const float* values;
  const int64_t* pix;
  int P;
  const int* runs;
  const ParView* views;
  mbir_par_geom g;
  void* sino;
  const void* w;
  const float* alpha;
  double* partials;
  hipStream_t s;
  bool bf16;      // sinogram / weights stored as bfloat16 (binned line search / update)
  float* adelta;  // line search: also store A(delta) (channel-major), or NULL
};

// Lane configurations: lanes per unit, rows per lane (VR = 4, 8 or 16:
// float4 loads, exactly when R % 4 == 0; the row arrays must then be 16-byte
// aligned), channels per unit.
struct LaneCfg {
  int lpc, vr, cps;
};
static inline bool fwd_vec(const mbir_par_geom& g) { return g.num_slices % 4 == 0; }
// The ray form's four-lane units evaluate each candidate's geometry once and
// share it over quad DPP moves; the rest of a candidate's cost is its VR x CPS
// FMAs per lane.  Measured at 720 views x 64 x 512 on a 512x512x64 grid:
// 4 lanes x 8 rows x 4 channels 4.2 ms; x 16 rows x 4 channels 4.4 ms; x 16
// rows x 2 channels 4.8 ms (5.4 ms before the geometry sharing); 8 x 8 x 4
// without sharing 6.8 ms.
static inline LaneCfg ray_lanes(int R, bool vec) {
  if (vec && R >= 24 && R % 8 == 0) return {4, 8, 4};
  if (vec) {
    const int lpc = R <= 8 ? 2 : (R <= 16 ? 4 : (R <= 32 ? 8 : 16));
    return {lpc, 4, lpc >= 8 ? 4 : (lpc == 4 ? 2 : 1)};
  }
  return {R <= 16 ? 16 : 64, 1, 4};
}
// The binned form reads a candidate's tap weights from LDS and loads its
// cylinder once per unit, so its per-candidate work shrinks with fewer lanes
// per unit.  Measured at 400 views x 512 channels, 1602- and 12814-pixel
// subsets: 32 rows as 4 lanes x 8 rows x 2 channels beats 8 x 4 x 4 by
// 1.05-1.55x; 64 rows keep 8 x 8 x 4.
static inline LaneCfg bin_lanes(int R, bool vec) {
  if (vec) {
    if (R <= 8) return {2, 4, 2};
    if (R <= 16) return {4, 4, 2};
    if (R <= 32) return R % 8 == 0 ? LaneCfg{4, 8, 2} : LaneCfg{8, 4, 4};
    return R % 8 == 0 ? LaneCfg{8, 8, 4} : LaneCfg{16, 4, 4};
  }
  return {R <= 16 ? 16 : 64, 1, 4};
}

// Workgroup count of the ray (ray = true) or binned form: the number of
// line-search partial pairs it writes.
static long fwd_blocks(const mbir_par_geom& g, bool ray) {
  const bool vec = fwd_vec(g);
  const LaneCfg c = ray ? ray_lanes(g.num_slices, vec) : bin_lanes(g.num_slices, vec);
  const long rows = (long)c.lpc * c.vr;
  const long z = (g.num_slices + rows - 1) / rows;
  const long cb = (long)(kBlock / c.lpc) * c.cps;
  return (long)g.num_views * ((g.num_channels + cb - 1) / cb) * z;
}

template <int LPC, int VR, int CPS, bool TWO, int MODE, bool WEIGHTED>
static void launch_ray(const FwdArgs& a) {
  const mbir_par_geom& g = a.g;
  constexpr int NB = 4;  // candidates per batch (one per lane of a four-lane unit)
  constexpr int CHW = (kBlock / LPC) * CPS;
  dim3 grid((g.num_channels + CHW - 1) / CHW, g.num_views,
            (g.num_slices + LPC * VR - 1) / (LPC * VR));
  par_fwd_ray_kernel<LPC, VR, CPS, NB, TWO, MODE, WEIGHTED><<<grid, kBlock, 0, a.s>>>(
      a.values, a.runs, a.views, g, static_cast<float*>(a.sino),
      static_cast<const float*>(a.w), a.alpha, a.partials, a.adelta);
}

template <int LPC, int VR, int CPS, bool TWO, int MODE, bool WEIGHTED, typename TS>
static void launch_bin(const FwdArgs& a) {
  const mbir_par_geom& g = a.g;
  constexpr int NB = 2;  // candidates per batch
  constexpr int CHW = (kBlock / LPC) * CPS;
  dim3 grid(g.num_views, (g.num_channels + CHW - 1) / CHW,
            (g.num_slices + LPC * VR - 1) / (LPC * VR));
  par_fwd_bin_kernel<LPC, VR, CPS, NB, TWO, MODE, WEIGHTED, TS><<<grid, kBlock, 0, a.s>>>(
      a.values, a.pix, a.P, a.views, g, static_cast<TS*>(a.sino),
      static_cast<const TS*>(a.w), a.alpha, a.partials, a.adelta);
}

// Mode dispatch for one lane configuration; bfloat16 storage serves the
// binned line search and update only.
template <int LPC, int VR, int CPS, bool TWO, bool RAY>
static int launch_fwd_mode(const FwdArgs& a) {
  if constexpr (!RAY) {
    if (a.bf16) {
      if (a.mode == FWD_LINESEARCH) {
        if (a.w) launch_bin<LPC, VR, CPS, TWO, FWD_LINESEARCH, true, bf16_bits>(a);
        else launch_bin<LPC, VR, CPS, TWO, FWD_LINESEARCH, false, bf16_bits>(a);
      } else {
        launch_bin<LPC, VR, CPS, TWO, FWD_UPDATE, false, bf16_bits>(a);
      }
      MBIR_CHECK_LAUNCH();
      return MBIR_OK;
    }
  }
#define MBIR_FWD_LAUNCH(MODE, W)                                \
  do {                                                          \
    if constexpr (RAY) launch_ray<LPC, VR, CPS, TWO, MODE, W>(a); \
    else launch_bin<LPC, VR, CPS, TWO, MODE, W, float>(a);      \
  } while (0)
  switch (a.mode) {
    case FWD_WRITE_VCR: MBIR_FWD_LAUNCH(FWD_WRITE_VCR, false); break;
    case FWD_WRITE_VRC: MBIR_FWD_LAUNCH(FWD_WRITE_VRC, false); break;
    case FWD_LINESEARCH:
      if (a.w) MBIR_FWD_LAUNCH(FWD_LINESEARCH, true);
      else MBIR_FWD_LAUNCH(FWD_LINESEARCH, false);
      break;
    default: MBIR_FWD_LAUNCH(FWD_UPDATE, false);
  }
#undef MBIR_FWD_LAUNCH
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

template <bool TWO>
static int launch_fwd_lanes(const FwdArgs& a, bool vec) {
  if (a.runs) {
    const LaneCfg c = ray_lanes(a.g.num_slices, vec);
    if (c.vr == 8) return launch_fwd_mode<4, 8, 4, TWO, true>(a);
    if (c.vr == 4) {
      switch (c.lpc) {
        case 2: return launch_fwd_mode<2, 4, 1, TWO, true>(a);
        case 4: return launch_fwd_mode<4, 4, 2, TWO, true>(a);
        case 8: return launch_fwd_mode<8, 4, 4, TWO, true>(a);
        default: return launch_fwd_mode<16, 4, 4, TWO, true>(a);
      }
    }
    return c.lpc == 16 ? launch_fwd_mode<16, 1, 4, TWO, true>(a)
                       : launch_fwd_mode<64, 1, 4, TWO, true>(a);
  }
  const LaneCfg c = bin_lanes(a.g.num_slices, vec);
  if (c.vr == 8) {
    return c.lpc == 4 ? launch_fwd_mode<4, 8, 2, TWO, false>(a)
                      : launch_fwd_mode<8, 8, 4, TWO, false>(a);
  }
  if (c.vr == 4) {
    switch (c.lpc) {
      case 2: return launch_fwd_mode<2, 4, 2, TWO, false>(a);
      case 4: return launch_fwd_mode<4, 4, 2, TWO, false>(a);
      case 8: return launch_fwd_mode<8, 4, 4, TWO, false>(a);
      default: return launch_fwd_mode<16, 4, 4, TWO, false>(a);
    }
  }
  return c.lpc == 16 ? launch_fwd_mode<16, 1, 4, TWO, false>(a)
                     : launch_fwd_mode<64, 1, 4, TWO, false>(a);
}

static inline bool aligned16(const void* p) {
  return (reinterpret_cast<uintptr_t>(p) & 15) == 0;
}

static int launch_fwd(const FwdArgs& a) {
  const bool vec = fwd_vec(a.g);
  // bfloat16 rows are read 4 at a time (8 bytes); 16-byte alignment covers it.
  if (vec && !(aligned16(a.values) && aligned16(a.sino) && (!a.w || aligned16(a.w))))
    return MBIR_ERR_ARG;
  if (a.bf16 && (a.runs || a.mode == FWD_WRITE_VCR || a.mode == FWD_WRITE_VRC))
    return MBIR_ERR_ARG;
  return a.g.two_tap ? launch_fwd_lanes<true>(a, vec) : launch_fwd_lanes<false>(a, vec);
}

}  // namespace mbir

using namespace mbir;

extern "C" {

int mbir_version(void) { return 3; }

int32_t mbir_par_back_splits(const mbir_par_geom* g, int32_t P) {
  if (!g || P <= 0) return 1;
  int pwg, rows;
  back_shape(*g, pwg, rows);
  const long wg = (long)((P + pwg - 1) / pwg) * ((g->num_slices + rows - 1) / rows);
  // Aim for >= 2048 workgroups (8 per CU) while keeping >= 8 views per split.
  long ns = (2048 + wg - 1) / wg;
  long max_ns = g->num_views / 8;
  if (ns > max_ns) ns = max_ns;
  if (ns < 1) ns = 1;
  return (int32_t)ns;
}

static int par_back_any(const void* sino, const void* w, const int64_t* pix, int32_t P,
                        const float* view_table, const mbir_par_geom* g, int32_t cp, int32_t ns,
                        float* out, mbir_stream_t stream, bool bf16) {
  int st = check_geom(g);
  if (st) return st;
  if (!sino || !pix || !view_table || !out || P < 0) return MBIR_ERR_ARG;
  if (cp != 1 && cp != 2) return MBIR_ERR_ARG;
  if (ns < 1 || ns > g->num_views) return MBIR_ERR_ARG;
  if (P == 0) return MBIR_OK;
  hipStream_t s = (hipStream_t)stream;
  const ParView* views = reinterpret_cast<const ParView*>(view_table);
  if (g->two_tap) return launch_back_nt<2>(sino, w, pix, P, views, *g, cp, ns, out, s, bf16);
  switch (2 * g->psf_radius + 1) {
    case 1: return launch_back_nt<1>(sino, w, pix, P, views, *g, cp, ns, out, s, bf16);
    case 3: return launch_back_nt<3>(sino, w, pix, P, views, *g, cp, ns, out, s, bf16);
    case 5: return launch_back_nt<5>(sino, w, pix, P, views, *g, cp, ns, out, s, bf16);
    case 7: return launch_back_nt<7>(sino, w, pix, P, views, *g, cp, ns, out, s, bf16);
    default: return launch_back_nt<9>(sino, w, pix, P, views, *g, cp, ns, out, s, bf16);
  }
}

int mbir_par_back(const float* sino, const float* w, const int64_t* pix, int32_t P,
                  const float* view_table, const mbir_par_geom* g, int32_t cp,
                  int32_t ns, float* out, mbir_stream_t stream) {
  return par_back_any(sino, w, pix, P, view_table, g, cp, ns, out, stream, false);
}

int mbir_par_back_bf16(const uint16_t* sino, const uint16_t* w, const int64_t* pix, int32_t P,
                       const float* view_table, const mbir_par_geom* g, int32_t ns,
                       float* out, mbir_stream_t stream) {
  return par_back_any(sino, w, pix, P, view_table, g, 1, ns, out, stream, true);
}

int mbir_reduce_splits(const float* parts, int32_t ns, int64_t n, float* out,
                       mbir_stream_t stream) {
  if (!parts || !out || ns < 1 || n < 0) return MBIR_ERR_ARG;
  if (n == 0) return MBIR_OK;
  int blocks = (int)((n + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  reduce_splits_kernel<<<blocks, 256, 0, (hipStream_t)stream>>>(parts, ns, n, out);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int mbir_par_forward(const float* values, const int64_t* pix, int32_t P, const int32_t* runs,
                     const float* view_table, const mbir_par_geom* g, int32_t layout,
                     float* out, mbir_stream_t stream) {
  int st = check_geom(g);
  if (st) return st;
  if (!values || !view_table || !out || P < 0 || (!pix && !runs)) return MBIR_ERR_ARG;
  if (layout != MBIR_LAYOUT_VRC && layout != MBIR_LAYOUT_VCR) return MBIR_ERR_ARG;
  FwdArgs a{layout == MBIR_LAYOUT_VRC ? FWD_WRITE_VRC : FWD_WRITE_VCR,
            values, pix, P, runs, reinterpret_cast<const ParView*>(view_table), *g, out,
            nullptr, nullptr, nullptr, (hipStream_t)stream, false, nullptr};
  return launch_fwd(a);
}

int32_t mbir_par_fwd_partials_count(const mbir_par_geom* g, int32_t ray) {
  if (!g || check_geom(g)) return 0;
  return (int32_t)fwd_blocks(*g, ray != 0);
}

int mbir_par_fwd_linesearch(const float* delta, const int64_t* pix, int32_t P,
                            const int32_t* runs, const float* view_table,
                            const mbir_par_geom* g, const float* error_vcr, const float* w,
                            float* adelta_vcr, double* partials, mbir_stream_t stream) {
  int st = check_geom(g);
  if (st) return st;
  if (!delta || !view_table || !partials || P < 0 || (!pix && !runs)) return MBIR_ERR_ARG;
  FwdArgs a{FWD_LINESEARCH, delta, pix, P, runs,
            reinterpret_cast<const ParView*>(view_table), *g, const_cast<float*>(error_vcr), w,
            nullptr, partials, (hipStream_t)stream, false, adelta_vcr};
  return launch_fwd(a);
}

int mbir_par_fwd_update(const float* delta, const int64_t* pix, int32_t P, const int32_t* runs,
                        const float* view_table, const mbir_par_geom* g, const float* alpha,
                        float* error_vcr, mbir_stream_t stream) {
  int st = check_geom(g);
  if (st) return st;
  if (!delta || !view_table || !error_vcr || !alpha || P < 0 || (!pix && !runs))
    return MBIR_ERR_ARG;
  FwdArgs a{FWD_UPDATE, delta, pix, P, runs, reinterpret_cast<const ParView*>(view_table), *g,
            error_vcr, nullptr, alpha, nullptr, (hipStream_t)stream, false, nullptr};
  return launch_fwd(a);
}

int mbir_par_fwd_linesearch_bf16(const float* delta, const int64_t* pix, int32_t P,
                                 const float* view_table, const mbir_par_geom* g,
                                 const uint16_t* error_vcr, const uint16_t* w, float* adelta_vcr,
                                 double* partials, mbir_stream_t stream) {
  int st = check_geom(g);
  if (st) return st;
  if (!delta || !pix || !view_table || !partials || P < 0) return MBIR_ERR_ARG;
  FwdArgs a{FWD_LINESEARCH, delta, pix, P, nullptr,
            reinterpret_cast<const ParView*>(view_table), *g, const_cast<uint16_t*>(error_vcr),
            w, nullptr, partials, (hipStream_t)stream, true, adelta_vcr};
  return launch_fwd(a);
}

int mbir_par_fwd_update_bf16(const float* delta, const int64_t* pix, int32_t P,
                             const float* view_table, const mbir_par_geom* g, const float* alpha,
                             uint16_t* error_vcr, mbir_stream_t stream) {
  int st = check_geom(g);
  if (st) return st;
  if (!delta || !pix || !view_table || !error_vcr || !alpha || P < 0) return MBIR_ERR_ARG;
  FwdArgs a{FWD_UPDATE, delta, pix, P, nullptr, reinterpret_cast<const ParView*>(view_table),
            *g, error_vcr, nullptr, alpha, nullptr, (hipStream_t)stream, true, nullptr};
  return launch_fwd(a);
}

}  // extern "C"
