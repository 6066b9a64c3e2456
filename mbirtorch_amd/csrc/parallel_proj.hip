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
// otherwise the reference's 2*psf+1 taps around round(n_p) are.
//
// Back projector (owner computes, atomic-free, deterministic): a workgroup
// owns a set of voxel cylinders and a view range.  Geometry for
// (view batch x pixels) is computed once into LDS; lanes of a wavefront then
// map to (pixel slot, detector row) so each tap reads a contiguous run of the
// channel-major sinogram row e[v, n, 0:R].  Small pixel sets split the view
// range over workgroups and leave partial sums reduced in a fixed order.
//
// Forward projector (per-view tile): a workgroup owns one view and a band of
// RC detector rows and keeps the (C x RC) sinogram tile in LDS; every pixel
// of the list adds its taps into the tile with LDS float atomics.  The tile
// is then written out, or -- for the VCD subset step -- reduced against the
// error sinogram (line search, FP64) or subtracted from it (update), so the
// subset direction's projection never goes to HBM.
#include "common.h"

namespace mbir {

constexpr int kBlock = 256;
constexpr int kMaxTaps = 9;  // psf_radius <= 4

// Geometry for one (view, pixel): tap base index and NT weights (zero for
// taps off the detector, base clamped so every tap index is in range when
// the weight is nonzero).
template <int NT>
__device__ __forceinline__ void par_taps(int64_t pix, const ParView& vw,
                                         const mbir_par_geom& g, int& base,
                                         float (&a)[NT]) {
  const int row = (int)(pix / g.num_cols);
  const int col = (int)(pix - (int64_t)row * g.num_cols);
  const float y_t = g.delta_voxel_row * ((float)row - g.row_center);
  const float x_t = g.delta_voxel * ((float)col - g.col_center);
  const float x = vw.cosv * x_t - vw.sinv * y_t;
  float t = x + g.det_channel_offset;
  if (!g.unit_channel) t = t / g.delta_det_channel;
  const float n_p = t + g.det_center;
  if (NT == 2) {
    base = (int)floorf(n_p);
  } else {
    base = (int)rintf(n_p) - g.psf_radius;
  }
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    const int n = base + k;
    float A = fmaxf(vw.halfw - fabsf(n_p - (float)n), 0.0f);
    A = fminf(A, vw.lmax) * vw.ws;
    a[k] = (n >= 0 && n < g.num_channels) ? A : 0.0f;
  }
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
// Forward projection into a per-view LDS tile.
enum FwdMode { FWD_WRITE_VCR = 0, FWD_WRITE_VRC = 1, FWD_LINESEARCH = 2, FWD_UPDATE = 3 };

template <int RC, int NT, int MODE, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void par_fwd_tile_kernel(
    const float* __restrict__ values, const int64_t* __restrict__ pix, int n_pix,
    const ParView* __restrict__ views, mbir_par_geom g, float* __restrict__ sino,
    const float* __restrict__ wts, const float* __restrict__ alpha,
    double* __restrict__ partials) {
  constexpr int TS = RC + 1;  // odd stride: conflict-free transposed reads
  constexpr int SLOTS = MBIR_WAVE / RC;
  extern __shared__ float smem[];
  const int C = g.num_channels, R = g.num_slices;
  float* tile = smem;
  int* s_n0 = reinterpret_cast<int*>(tile + (size_t)C * TS);
  float* s_a = reinterpret_cast<float*>(s_n0 + kBlock);

  const int v = blockIdx.x, r0 = blockIdx.y * RC;
  const int lane = threadIdx.x & (MBIR_WAVE - 1), wave = threadIdx.x / MBIR_WAVE;
  const int slot = lane / RC, rl = lane % RC;
  const bool r_ok = (r0 + rl) < R;
  const ParView vw = views[v];

  for (int i = threadIdx.x; i < C * TS; i += kBlock) tile[i] = 0.0f;

  for (int pc = 0; pc < n_pix; pc += kBlock) {
    __syncthreads();
    {
      const int pg = pc + threadIdx.x;
      int base = 0;
      float a[NT];
      if (pg < n_pix) {
        par_taps<NT>(pix[pg], vw, g, base, a);
      } else {
#pragma unroll
        for (int k = 0; k < NT; ++k) a[k] = 0.0f;
      }
      s_n0[threadIdx.x] = base;
#pragma unroll
      for (int k = 0; k < NT; ++k) s_a[k * kBlock + threadIdx.x] = a[k];
    }
    __syncthreads();
    const int cnt = min(kBlock, n_pix - pc);
    if (r_ok) {
      for (int jj = wave * MBIR_WAVE; jj < wave * MBIR_WAVE + MBIR_WAVE; jj += SLOTS) {
        const int j = jj + slot;
        if (j < cnt) {
          const float xv = values[(size_t)(pc + j) * R + r0 + rl];
          const int base = s_n0[j];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const float a = s_a[t * kBlock + j];
            if (a != 0.0f) atomicAdd(&tile[(base + t) * TS + rl], a * xv);
          }
        }
      }
    }
  }
  __syncthreads();

  if (MODE == FWD_WRITE_VRC) {
    // out[v, r, c]: consecutive threads take consecutive channels.
    for (int i = threadIdx.x; i < C * RC; i += kBlock) {
      const int rr = i / C, c = i - rr * C;
      if (r0 + rr < R) sino[((size_t)v * R + r0 + rr) * C + c] = tile[c * TS + rr];
    }
  } else {
    double lin = 0.0, quad = 0.0;
    float al = 0.0f;
    if (MODE == FWD_UPDATE) al = alpha[0];
    for (int i = threadIdx.x; i < C * RC; i += kBlock) {
      const int c = i / RC, rr = i - c * RC;
      if (r0 + rr >= R) continue;
      const size_t o = ((size_t)v * C + c) * R + r0 + rr;
      const float t = tile[c * TS + rr];
      if (MODE == FWD_WRITE_VCR) {
        sino[o] = t;
      } else if (MODE == FWD_LINESEARCH) {
        const float e = sino[o];
        if (WEIGHTED) {
          const float w = wts[o];
          lin += (double)((w * e) * t);
          quad += (double)((t * t) * w);
        } else {
          lin += (double)(e * t);
          quad += (double)(t * t);
        }
      } else {  // FWD_UPDATE
        sino[o] = sino[o] - al * t;
      }
    }
    if (MODE == FWD_LINESEARCH) {
      block_sum2_f64(lin, quad);
      if (threadIdx.x == 0) {
        const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        partials[2 * b] = lin;
        partials[2 * b + 1] = quad;
      }
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
  if (!(g->delta_det_channel > 0.0f)) return MBIR_ERR_GEOMETRY;
  return MBIR_OK;
}

static inline int back_rw(int R) { return R >= 64 ? 64 : next_pow2(R); }

static inline int back_pix_per_wg(int RW) {
  return (kBlock / MBIR_WAVE) * (MBIR_WAVE / RW) * (RW >= 4 ? 8 : 4);
}

// Rows per forward tile: largest power of two <= 64 that covers R and keeps
// the tile within 60 KiB of LDS.
static inline int fwd_rc(const mbir_par_geom* g) {
  int rc = g->num_slices >= 64 ? 64 : next_pow2(g->num_slices);
  while (rc > 1 && (size_t)g->num_channels * (rc + 1) * 4 + kBlock * 4 * (1 + kMaxTaps) >
                       60 * 1024)
    rc >>= 1;
  return rc;
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

template <int RC, int NT>
static int launch_fwd_rc(int mode, const float* values, const int64_t* pix, int P,
                         const ParView* views, const mbir_par_geom& g, float* sino,
                         const float* w, const float* alpha, double* partials,
                         hipStream_t s) {
  dim3 grid(g.num_views, (g.num_slices + RC - 1) / RC);
  const size_t lds = (size_t)g.num_channels * (RC + 1) * 4 + kBlock * 4 * (1 + NT);
  switch (mode) {
    case FWD_WRITE_VCR:
      par_fwd_tile_kernel<RC, NT, FWD_WRITE_VCR, false><<<grid, kBlock, lds, s>>>(
          values, pix, P, views, g, sino, w, alpha, partials);
      break;
    case FWD_WRITE_VRC:
      par_fwd_tile_kernel<RC, NT, FWD_WRITE_VRC, false><<<grid, kBlock, lds, s>>>(
          values, pix, P, views, g, sino, w, alpha, partials);
      break;
    case FWD_LINESEARCH:
      if (w)
        par_fwd_tile_kernel<RC, NT, FWD_LINESEARCH, true><<<grid, kBlock, lds, s>>>(
            values, pix, P, views, g, sino, w, alpha, partials);
      else
        par_fwd_tile_kernel<RC, NT, FWD_LINESEARCH, false><<<grid, kBlock, lds, s>>>(
            values, pix, P, views, g, sino, w, alpha, partials);
      break;
    default:
      par_fwd_tile_kernel<RC, NT, FWD_UPDATE, false><<<grid, kBlock, lds, s>>>(
          values, pix, P, views, g, sino, w, alpha, partials);
      break;
  }
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

template <int NT>
static int launch_fwd_nt(int mode, const float* values, const int64_t* pix, int P,
                         const ParView* views, const mbir_par_geom& g, float* sino,
                         const float* w, const float* alpha, double* partials,
                         hipStream_t s) {
  switch (fwd_rc(&g)) {
    case 64: return launch_fwd_rc<64, NT>(mode, values, pix, P, views, g, sino, w, alpha, partials, s);
    case 32: return launch_fwd_rc<32, NT>(mode, values, pix, P, views, g, sino, w, alpha, partials, s);
    case 16: return launch_fwd_rc<16, NT>(mode, values, pix, P, views, g, sino, w, alpha, partials, s);
    case 8: return launch_fwd_rc<8, NT>(mode, values, pix, P, views, g, sino, w, alpha, partials, s);
    case 4: return launch_fwd_rc<4, NT>(mode, values, pix, P, views, g, sino, w, alpha, partials, s);
    case 2: return launch_fwd_rc<2, NT>(mode, values, pix, P, views, g, sino, w, alpha, partials, s);
    default: return launch_fwd_rc<1, NT>(mode, values, pix, P, views, g, sino, w, alpha, partials, s);
  }
}

static int launch_fwd(int mode, const float* values, const int64_t* pix, int P,
                      const ParView* views, const mbir_par_geom& g, float* sino,
                      const float* w, const float* alpha, double* partials, hipStream_t s) {
  if (g.two_tap) return launch_fwd_nt<2>(mode, values, pix, P, views, g, sino, w, alpha, partials, s);
  switch (2 * g.psf_radius + 1) {
    case 1: return launch_fwd_nt<1>(mode, values, pix, P, views, g, sino, w, alpha, partials, s);
    case 3: return launch_fwd_nt<3>(mode, values, pix, P, views, g, sino, w, alpha, partials, s);
    case 5: return launch_fwd_nt<5>(mode, values, pix, P, views, g, sino, w, alpha, partials, s);
    case 7: return launch_fwd_nt<7>(mode, values, pix, P, views, g, sino, w, alpha, partials, s);
    default: return launch_fwd_nt<9>(mode, values, pix, P, views, g, sino, w, alpha, partials, s);
  }
}

}  // namespace mbir

using namespace mbir;

extern "C" {

int mbir_version(void) { return 1; }

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

int mbir_par_forward(const float* values, const int64_t* pix, int32_t P,
                     const float* view_table, const mbir_par_geom* g, int32_t layout,
                     float* out, mbir_stream_t stream) {
  int st = check_geom(g);
  if (st) return st;
  if (!values || !pix || !view_table || !out || P < 0) return MBIR_ERR_ARG;
  if (layout != MBIR_LAYOUT_VRC && layout != MBIR_LAYOUT_VCR) return MBIR_ERR_ARG;
  const int mode = layout == MBIR_LAYOUT_VRC ? FWD_WRITE_VRC : FWD_WRITE_VCR;
  return launch_fwd(mode, values, pix, P, reinterpret_cast<const ParView*>(view_table), *g,
                    out, nullptr, nullptr, nullptr, (hipStream_t)stream);
}

int32_t mbir_par_fwd_partials_count(const mbir_par_geom* g) {
  if (!g) return 0;
  const int rc = fwd_rc(g);
  return g->num_views * ((g->num_slices + rc - 1) / rc);
}

int mbir_par_fwd_linesearch(const float* delta, const int64_t* pix, int32_t P,
                            const float* view_table, const mbir_par_geom* g,
                            const float* error_vcr, const float* w, double* partials,
                            mbir_stream_t stream) {
  int st = check_geom(g);
  if (st) return st;
  if (!delta || !pix || !view_table || !error_vcr || !partials || P < 0) return MBIR_ERR_ARG;
  return launch_fwd(FWD_LINESEARCH, delta, pix, P,
                    reinterpret_cast<const ParView*>(view_table), *g,
                    const_cast<float*>(error_vcr), w, nullptr, partials, (hipStream_t)stream);
}

int mbir_par_fwd_update(const float* delta, const int64_t* pix, int32_t P,
                        const float* view_table, const mbir_par_geom* g, const float* alpha,
                        float* error_vcr, mbir_stream_t stream) {
  int st = check_geom(g);
  if (st) return st;
  if (!delta || !pix || !view_table || !error_vcr || !alpha || P < 0) return MBIR_ERR_ARG;
  return launch_fwd(FWD_UPDATE, delta, pix, P, reinterpret_cast<const ParView*>(view_table),
                    *g, error_vcr, nullptr, alpha, nullptr, (hipStream_t)stream);
}

}  // extern "C"
