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
//    view and 256 channels; chunks of the list are counting-sorted by channel
//    in LDS, each bin ordered by list position, and every thread gathers its
//    channel's bin.  Its epilogue writes the projection, or -- for the VCD
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
  const int row = (int)(pix / g.num_cols);
  const int col = (int)(pix - (int64_t)row * g.num_cols);
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
template <int RW, int NT, int CP, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void par_back_kernel(
    const float* __restrict__ sino, const float* __restrict__ wts,
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
  const int r = blockIdx.z * RW + rl;
  const bool r_ok = r < R;
  const int p_base = blockIdx.x * PIX_WG;
  const int v_begin = blockIdx.y * views_per_split;
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
        const float* sv = sino + vofs;
        const float* wv = WEIGHTED ? wts + vofs : nullptr;
#pragma unroll
        for (int k = 0; k < ACC; ++k) {
          const int q = b * PIX_WG + wave * PIX_WAVE + k * SLOTS + slot;
          const int base = s_n0[q];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const float a = s_a[t * VB * PIX_WG + q];
            const int n = min(max(base + t, 0), C - 1);
            float e = sv[(size_t)n * R];
            if (WEIGHTED) e = e * wv[(size_t)n * R];
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
      if (pg < n_pix) out[((size_t)blockIdx.y * n_pix + pg) * R + r] = acc[k];
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
// Shared forward helpers: accumulate RPT rows of one voxel cylinder.
template <int RPT, bool VEC>
__device__ __forceinline__ void acc_rows(float (&acc)[RPT], float A, const float* xp, int nr) {
  if (VEC) {
#pragma unroll
    for (int k = 0; k < RPT; k += 4) {
      if (k < nr) {
        const float4 x4 = *reinterpret_cast<const float4*>(xp + k);
        acc[k] = fmaf(A, x4.x, acc[k]);
        acc[k + 1] = fmaf(A, x4.y, acc[k + 1]);
        acc[k + 2] = fmaf(A, x4.z, acc[k + 2]);
        acc[k + 3] = fmaf(A, x4.w, acc[k + 3]);
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < RPT; ++k)
      if (k < nr) acc[k] = fmaf(A, xp[k], acc[k]);
  }
}

enum FwdMode { FWD_WRITE_VCR = 0, FWD_WRITE_VRC = 1, FWD_LINESEARCH = 2, FWD_UPDATE = 3 };

// Epilogue for a thread owning sinogram entries (v, n, r0 .. r0+nr).
template <int RPT, int MODE, bool WEIGHTED>
__device__ __forceinline__ void fwd_epilogue(const float (&acc)[RPT], int v, int n, int r0, int nr,
                                             const mbir_par_geom& g, float* __restrict__ sino,
                                             const float* __restrict__ wts, float al,
                                             double& lin, double& quad) {
  const int R = g.num_slices, C = g.num_channels;
  if (MODE == FWD_WRITE_VRC) {
#pragma unroll
    for (int k = 0; k < RPT; ++k)
      if (k < nr) sino[((size_t)v * R + r0 + k) * C + n] = acc[k];
    return;
  }
  const size_t o = ((size_t)v * C + n) * R + r0;
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    if (k >= nr) continue;
    const float t = acc[k];
    if (MODE == FWD_WRITE_VCR) {
      sino[o + k] = t;
    } else if (MODE == FWD_LINESEARCH) {
      const float e = sino[o + k];
      if (WEIGHTED) {
        const float w = wts[o + k];
        lin += (double)((w * e) * t);
        quad += (double)((t * t) * w);
      } else {
        lin += (double)(e * t);
        quad += (double)(t * t);
      }
    } else {
      sino[o + k] = sino[o + k] - al * t;
    }
  }
}

// ---------------------------------------------------------------------------
// Ray-driven forward for row-structured pixel lists.  runs[3*i .. 3*i+2] =
// (list position of the run's first pixel, first column, run length) for
// image row i; a length of 0 means the row holds no listed pixel.
constexpr int kRayBlock = 128;

template <int RPT, int NT, bool VEC, int MODE>
__global__ __launch_bounds__(kRayBlock) void par_fwd_ray_kernel(
    const float* __restrict__ values, const int* __restrict__ runs,
    const ParView* __restrict__ views, mbir_par_geom g, float* __restrict__ sino) {
  const int n = blockIdx.x * kRayBlock + threadIdx.x;
  const int v = blockIdx.y, r0 = blockIdx.z * RPT;
  const int R = g.num_slices, C = g.num_channels, Nr = g.num_rows, Nc = g.num_cols;
  if (n >= C) return;  // no barriers in this kernel
  const int nr = min(RPT, R - r0);
  const ParView vw = views[v];
  float acc[RPT];
#pragma unroll
  for (int k = 0; k < RPT; ++k) acc[k] = 0.0f;

  // Candidate windows come from the inverse map in float32 with a half-pixel
  // margin; every candidate is then checked with the exact forward chain.
  const float x_target = ((float)n - g.det_center) * g.delta_det_channel - g.det_channel_offset;
  const float cd = vw.cosv * g.delta_voxel, sd = vw.sinv * g.delta_voxel_row;
  if (fabsf(cd) >= fabsf(sd)) {
    const float hj = vw.halfw * g.delta_det_channel / fabsf(cd) + 0.5f;
    for (int i = 0; i < Nr; ++i) {
      const int start = runs[3 * i], jmin = runs[3 * i + 1], cnt = runs[3 * i + 2];
      if (cnt <= 0) continue;
      const float y_t = g.delta_voxel_row * ((float)i - g.row_center);
      const float jstar = (x_target + vw.sinv * y_t) / cd + g.col_center;
      const int jlo = max((int)floorf(jstar - hj), jmin);
      const int jhi = min((int)ceilf(jstar + hj), jmin + cnt - 1);
      for (int j = jlo; j <= jhi; ++j) {
        float A;
        if (par_tap_of<NT>(i, j, n, vw, g, A))
          acc_rows<RPT, VEC>(acc, A, values + (size_t)(start + j - jmin) * R + r0, nr);
      }
    }
  } else {
    const float hi = vw.halfw * g.delta_det_channel / fabsf(sd) + 0.5f;
    for (int j = 0; j < Nc; ++j) {
      const float x_t = g.delta_voxel * ((float)j - g.col_center);
      const float istar = (vw.cosv * x_t - x_target) / sd + g.row_center;
      const int ilo = max((int)floorf(istar - hi), 0);
      const int ihi = min((int)ceilf(istar + hi), Nr - 1);
      for (int i = ilo; i <= ihi; ++i) {
        const int start = runs[3 * i], jmin = runs[3 * i + 1], cnt = runs[3 * i + 2];
        if (j < jmin || j >= jmin + cnt) continue;
        float A;
        if (par_tap_of<NT>(i, j, n, vw, g, A))
          acc_rows<RPT, VEC>(acc, A, values + (size_t)(start + j - jmin) * R + r0, nr);
      }
    }
  }
  double lin = 0.0, quad = 0.0;
  fwd_epilogue<RPT, MODE, false>(acc, v, n, r0, nr, g, sino, nullptr, 0.0f, lin, quad);
}

// ---------------------------------------------------------------------------
// Binned forward for arbitrary pixel lists.  Workgroup = (view, 256 channels,
// RPT rows); thread t owns channel c0 + t.
template <int NT>
struct BinCfg {
  static constexpr int CH = NT <= 2 ? 1024 : (NT <= 3 ? 512 : 256);  // pixels per chunk
  static constexpr int PPT = CH / kBlock;                             // pixels per thread
  static constexpr int ENT = CH * NT;                                 // entry capacity
};

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

template <int RPT, int NT, bool VEC, int MODE, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void par_fwd_bin_kernel(
    const float* __restrict__ values, const int64_t* __restrict__ pix, int n_pix,
    const ParView* __restrict__ views, mbir_par_geom g, float* __restrict__ sino,
    const float* __restrict__ wts, const float* __restrict__ alpha,
    double* __restrict__ partials) {
  using Cfg = BinCfg<NT>;
  __shared__ int s_cnt[kBlock + 1];
  __shared__ int s_off[kBlock + 1];
  __shared__ int s_tmp[4];
  __shared__ int s_ej[Cfg::ENT];
  __shared__ float s_ea[Cfg::ENT];

  const int tid = threadIdx.x;
  const int v = blockIdx.x, c0 = blockIdx.y * kBlock, r0 = blockIdx.z * RPT;
  const int R = g.num_slices, C = g.num_channels;
  const int n = c0 + tid;
  const int nr = min(RPT, R - r0);
  const ParView vw = views[v];
  float acc[RPT];
#pragma unroll
  for (int k = 0; k < RPT; ++k) acc[k] = 0.0f;

  for (int pc = 0; pc < n_pix; pc += Cfg::CH) {
    // Geometry of this thread's pixels; taps landing in [c0, c0+256) count.
    int bin[Cfg::PPT][NT];
    float wgt[Cfg::PPT][NT];
    s_cnt[tid] = 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Cfg::PPT; ++q) {
      const int j = pc + tid + q * kBlock;
      int base = 0;
      float a[NT];
      if (j < n_pix) {
        par_taps<NT>(pix[j], vw, g, base, a);
      } else {
#pragma unroll
        for (int t = 0; t < NT; ++t) a[t] = 0.0f;
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int b = base + t - c0;
        const bool in = a[t] != 0.0f && b >= 0 && b < kBlock;
        bin[q][t] = in ? b : -1;
        wgt[q][t] = a[t];
        if (in) atomicAdd(&s_cnt[b], 1);
      }
    }
    __syncthreads();
    const int mine = s_cnt[tid];
    const int off = block_exclusive_scan(mine, s_tmp);
    s_off[tid] = off;
    if (tid == kBlock - 1) s_off[kBlock] = off + mine;
    s_cnt[tid] = off;  // scatter cursor
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Cfg::PPT; ++q) {
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int b = bin[q][t];
        if (b >= 0) {
          const int slot = atomicAdd(&s_cnt[b], 1);
          s_ej[slot] = tid + q * kBlock;  // position within the chunk
          s_ea[slot] = wgt[q][t];
        }
      }
    }
    __syncthreads();
    // Order this channel's bin by list position (bins are small for random
    // subsets), then gather.
    const int e0 = s_off[tid], e1 = s_off[tid + 1];
    for (int e = e0 + 1; e < e1; ++e) {
      const int kj = s_ej[e];
      const float ka = s_ea[e];
      int f = e - 1;
      while (f >= e0 && s_ej[f] > kj) {
        s_ej[f + 1] = s_ej[f];
        s_ea[f + 1] = s_ea[f];
        --f;
      }
      s_ej[f + 1] = kj;
      s_ea[f + 1] = ka;
    }
    if (n < C) {
      for (int e = e0; e < e1; ++e)
        acc_rows<RPT, VEC>(acc, s_ea[e], values + (size_t)(pc + s_ej[e]) * R + r0, nr);
    }
    __syncthreads();
  }

  double lin = 0.0, quad = 0.0;
  float al = 0.0f;
  if (MODE == FWD_UPDATE) al = alpha[0];
  if (n < C) fwd_epilogue<RPT, MODE, WEIGHTED>(acc, v, n, r0, nr, g, sino, wts, al, lin, quad);
  if (MODE == FWD_LINESEARCH) {
    block_sum2_f64(lin, quad);
    if (tid == 0) {
      const size_t b = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
      partials[2 * b] = lin;
      partials[2 * b + 1] = quad;
    }
  }
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

// Rows per forward thread: 64, or the smallest power of two >= 4 covering R.
static inline int fwd_rpt(int R) {
  if (R > 32) return 64;
  if (R > 16) return 32;
  if (R > 8) return 16;
  if (R > 4) return 8;
  return 4;
}

template <int RW, int NT>
static int launch_back_rw(const float* sino, const float* w, const int64_t* pix, int P,
                          const ParView* views, const mbir_par_geom& g, int cp, int ns,
                          float* out, hipStream_t s) {
  const int pwg = back_pix_per_wg(RW);
  const int vps = (g.num_views + ns - 1) / ns;
  dim3 grid((P + pwg - 1) / pwg, ns, (g.num_slices + RW - 1) / RW);
  if (cp == 1) {
    if (w) par_back_kernel<RW, NT, 1, true><<<grid, kBlock, 0, s>>>(sino, w, pix, P, views, g, vps, out);
    else par_back_kernel<RW, NT, 1, false><<<grid, kBlock, 0, s>>>(sino, w, pix, P, views, g, vps, out);
  } else {
    if (w) par_back_kernel<RW, NT, 2, true><<<grid, kBlock, 0, s>>>(sino, w, pix, P, views, g, vps, out);
    else par_back_kernel<RW, NT, 2, false><<<grid, kBlock, 0, s>>>(sino, w, pix, P, views, g, vps, out);
  }
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

template <int NT>
static int launch_back_nt(const float* sino, const float* w, const int64_t* pix, int P,
                          const ParView* views, const mbir_par_geom& g, int cp, int ns,
                          float* out, hipStream_t s) {
  switch (back_rw(g.num_slices)) {
    case 64: return launch_back_rw<64, NT>(sino, w, pix, P, views, g, cp, ns, out, s);
    case 32: return launch_back_rw<32, NT>(sino, w, pix, P, views, g, cp, ns, out, s);
    case 16: return launch_back_rw<16, NT>(sino, w, pix, P, views, g, cp, ns, out, s);
    case 8: return launch_back_rw<8, NT>(sino, w, pix, P, views, g, cp, ns, out, s);
    case 4: return launch_back_rw<4, NT>(sino, w, pix, P, views, g, cp, ns, out, s);
    case 2: return launch_back_rw<2, NT>(sino, w, pix, P, views, g, cp, ns, out, s);
    default: return launch_back_rw<1, NT>(sino, w, pix, P, views, g, cp, ns, out, s);
  }
}

// Forward dispatch: ray kernel when `runs` is given (write modes), binned
// kernel otherwise.
struct FwdArgs {
  int mode;
  const float* values;
  const int64_t* pix;
  int P;
  const int* runs;
  const ParView* views;
  mbir_par_geom g;
  float* sino;
  const float* w;
  const float* alpha;
  double* partials;
  hipStream_t s;
};

template <int RPT, int NT, bool VEC>
static int launch_fwd3(const FwdArgs& a) {
  const mbir_par_geom& g = a.g;
  if (a.runs) {
    dim3 grid((g.num_channels + kRayBlock - 1) / kRayBlock, g.num_views,
              (g.num_slices + RPT - 1) / RPT);
    if (a.mode == FWD_WRITE_VRC)
      par_fwd_ray_kernel<RPT, NT, VEC, FWD_WRITE_VRC><<<grid, kRayBlock, 0, a.s>>>(
          a.values, a.runs, a.views, g, a.sino);
    else
      par_fwd_ray_kernel<RPT, NT, VEC, FWD_WRITE_VCR><<<grid, kRayBlock, 0, a.s>>>(
          a.values, a.runs, a.views, g, a.sino);
  } else {
    dim3 grid(g.num_views, (g.num_channels + kBlock - 1) / kBlock,
              (g.num_slices + RPT - 1) / RPT);
    switch (a.mode) {
      case FWD_WRITE_VCR:
        par_fwd_bin_kernel<RPT, NT, VEC, FWD_WRITE_VCR, false><<<grid, kBlock, 0, a.s>>>(
            a.values, a.pix, a.P, a.views, g, a.sino, a.w, a.alpha, a.partials);
        break;
      case FWD_WRITE_VRC:
        par_fwd_bin_kernel<RPT, NT, VEC, FWD_WRITE_VRC, false><<<grid, kBlock, 0, a.s>>>(
            a.values, a.pix, a.P, a.views, g, a.sino, a.w, a.alpha, a.partials);
        break;
      case FWD_LINESEARCH:
        if (a.w)
          par_fwd_bin_kernel<RPT, NT, VEC, FWD_LINESEARCH, true><<<grid, kBlock, 0, a.s>>>(
              a.values, a.pix, a.P, a.views, g, a.sino, a.w, a.alpha, a.partials);
        else
          par_fwd_bin_kernel<RPT, NT, VEC, FWD_LINESEARCH, false><<<grid, kBlock, 0, a.s>>>(
              a.values, a.pix, a.P, a.views, g, a.sino, a.w, a.alpha, a.partials);
        break;
      default:
        par_fwd_bin_kernel<RPT, NT, VEC, FWD_UPDATE, false><<<grid, kBlock, 0, a.s>>>(
            a.values, a.pix, a.P, a.views, g, a.sino, a.w, a.alpha, a.partials);
    }
  }
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

template <int RPT, int NT>
static int launch_fwd2(const FwdArgs& a) {
  // float4 rows need R % 4 == 0 (row chunks start at multiples of RPT >= 4)
  // and a 16-byte aligned value array.
  if (a.g.num_slices % 4 == 0 && (reinterpret_cast<uintptr_t>(a.values) & 15) == 0)
    return launch_fwd3<RPT, NT, true>(a);
  return launch_fwd3<RPT, NT, false>(a);
}

template <int NT>
static int launch_fwd1(const FwdArgs& a) {
  switch (fwd_rpt(a.g.num_slices)) {
    case 64: return launch_fwd2<64, NT>(a);
    case 32: return launch_fwd2<32, NT>(a);
    case 16: return launch_fwd2<16, NT>(a);
    case 8: return launch_fwd2<8, NT>(a);
    default: return launch_fwd2<4, NT>(a);
  }
}

static int launch_fwd(const FwdArgs& a) {
  if (a.g.two_tap) return launch_fwd1<2>(a);
  switch (2 * a.g.psf_radius + 1) {
    case 1: return launch_fwd1<1>(a);
    case 3: return launch_fwd1<3>(a);
    case 5: return launch_fwd1<5>(a);
    case 7: return launch_fwd1<7>(a);
    default: return launch_fwd1<9>(a);
  }
}

}  // namespace mbir

using namespace mbir;

extern "C" {

int mbir_version(void) { return 2; }

int32_t mbir_par_back_splits(const mbir_par_geom* g, int32_t P) {
  if (!g || P <= 0) return 1;
  const int RW = back_rw(g->num_slices);
  const int pwg = back_pix_per_wg(RW);
  const long wg = (long)((P + pwg - 1) / pwg) * ((g->num_slices + RW - 1) / RW);
  // Aim for >= 2048 workgroups (8 per CU) while keeping >= 8 views per split.
  long ns = (2048 + wg - 1) / wg;
  long max_ns = g->num_views / 8;
  if (ns > max_ns) ns = max_ns;
  if (ns < 1) ns = 1;
  return (int32_t)ns;
}

int mbir_par_back(const float* sino, const float* w, const int64_t* pix, int32_t P,
                  const float* view_table, const mbir_par_geom* g, int32_t cp,
                  int32_t ns, float* out, mbir_stream_t stream) {
  int st = check_geom(g);
  if (st) return st;
  if (!sino || !pix || !view_table || !out || P < 0) return MBIR_ERR_ARG;
  if (cp != 1 && cp != 2) return MBIR_ERR_ARG;
  if (ns < 1 || ns > g->num_views) return MBIR_ERR_ARG;
  if (P == 0) return MBIR_OK;
  hipStream_t s = (hipStream_t)stream;
  const ParView* views = reinterpret_cast<const ParView*>(view_table);
  if (g->two_tap) return launch_back_nt<2>(sino, w, pix, P, views, *g, cp, ns, out, s);
  switch (2 * g->psf_radius + 1) {
    case 1: return launch_back_nt<1>(sino, w, pix, P, views, *g, cp, ns, out, s);
    case 3: return launch_back_nt<3>(sino, w, pix, P, views, *g, cp, ns, out, s);
    case 5: return launch_back_nt<5>(sino, w, pix, P, views, *g, cp, ns, out, s);
    case 7: return launch_back_nt<7>(sino, w, pix, P, views, *g, cp, ns, out, s);
    default: return launch_back_nt<9>(sino, w, pix, P, views, *g, cp, ns, out, s);
  }
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
  if (!values || !pix || !view_table || !out || P < 0) return MBIR_ERR_ARG;
  if (layout != MBIR_LAYOUT_VRC && layout != MBIR_LAYOUT_VCR) return MBIR_ERR_ARG;
  FwdArgs a{layout == MBIR_LAYOUT_VRC ? FWD_WRITE_VRC : FWD_WRITE_VCR,
            values, pix, P, runs, reinterpret_cast<const ParView*>(view_table), *g, out,
            nullptr, nullptr, nullptr, (hipStream_t)stream};
  return launch_fwd(a);
}

int32_t mbir_par_fwd_partials_count(const mbir_par_geom* g) {
  if (!g) return 0;
  const int rpt = fwd_rpt(g->num_slices);
  return g->num_views * ((g->num_channels + kBlock - 1) / kBlock) *
         ((g->num_slices + rpt - 1) / rpt);
}

int mbir_par_fwd_linesearch(const float* delta, const int64_t* pix, int32_t P,
                            const float* view_table, const mbir_par_geom* g,
                            const float* error_vcr, const float* w, double* partials,
                            mbir_stream_t stream) {
  int st = check_geom(g);
  if (st) return st;
  if (!delta || !pix || !view_table || !error_vcr || !partials || P < 0) return MBIR_ERR_ARG;
  FwdArgs a{FWD_LINESEARCH, delta, pix, P, nullptr,
            reinterpret_cast<const ParView*>(view_table), *g, const_cast<float*>(error_vcr), w,
            nullptr, partials, (hipStream_t)stream};
  return launch_fwd(a);
}

int mbir_par_fwd_update(const float* delta, const int64_t* pix, int32_t P,
                        const float* view_table, const mbir_par_geom* g, const float* alpha,
                        float* error_vcr, mbir_stream_t stream) {
  int st = check_geom(g);
  if (st) return st;
  if (!delta || !pix || !view_table || !error_vcr || !alpha || P < 0) return MBIR_ERR_ARG;
  FwdArgs a{FWD_UPDATE, delta, pix, P, nullptr, reinterpret_cast<const ParView*>(view_table),
            *g, error_vcr, nullptr, alpha, nullptr, (hipStream_t)stream};
  return launch_fwd(a);
}

}  // extern "C"
