// Cone-beam projectors for MI355X (gfx950): flat or curved detector,
// circular or helical scans.
//
// Geometry for one (view, pixel), in the reference's float32 order
// (cone_beam.py:59-158): rotated (x, y), per-pixel magnification
// 1/(1/M - y/SDD), channel coordinate n_p (flat u = mag*x, curved
// u = SDD*atan2(x, SID - y)), projected width W_c, weight scale, and the
// vertical affine map m(k) = m0 + W_r*k from global slice k to detector row.
// The cone-angle factor enters as 1/cos(phi) = sqrt(1 + (mag*z_k/SDD)^2)
// (algebraically cos(atan2(mag*z_k, SDD))^-1).  Taps follow the reference's
// windows: channels round(n_p) +- psf, back rows round(m(k)) +- psf, forward
// slices round(k_m) +- bp_psf.
//
// Back (cone_beam.py:237-299): owner computes, atomic-free.  Geometry for a
// (view batch x pixel set) goes to LDS once; lanes then map to slices, so the
// taps of consecutive slices read consecutive detector rows of the
// channel-major sinogram e[v, n, m].
//
// Forward (cone_beam.py:161-234): an atomic-free binned gather -- a
// workgroup owns one view, 64 channels and a band of detector rows (one lane
// per row); the pixel list is counting-sorted by channel in LDS and each
// row gathers its voxel taps (detector-side vertical fan) per channel entry.
// Epilogues as in the parallel projector: write, FP64 line search, or error
// update.
#include "common.h"

namespace mbir {

constexpr int kCBlock = 256;
constexpr int kCMaxTaps = 9;

struct ConeView {
  float cosv, sinv, angle, z_off;
};

struct ConePixel {
  float n_p, Wc, ws, mag, m0, Wr;
};

__device__ __forceinline__ void cone_pixel(int64_t pix, const ConeView& vw,
                                           const mbir_cone_geom& g, ConePixel& o) {
  const int pj = (int)pix;  // pixel indices lie in [0, rows*cols) < 2^31
  const int row = pj / g.num_cols;
  const int col = pj - row * g.num_cols;
  const float y_t = g.delta_voxel_row * ((float)row - g.row_center);
  const float x_t = g.delta_voxel * ((float)col - g.col_center);
  const float x = vw.cosv * x_t - vw.sinv * y_t;
  const float y = vw.sinv * x_t + vw.cosv * y_t;
  const float sdd = g.source_detector_dist;
  const float mag = 1.0f / (g.inv_magnification - y / sdd);
  float u, theta;
  if (!g.curved) {
    u = mag * x;
    theta = atan2f(u, sdd);
  } else {
    u = sdd * atan2f(x, g.source_iso_dist - y);
    theta = u / sdd;
  }
  float t = u + g.det_channel_offset;
  if (!g.unit_channel) t = t / g.delta_det_channel;
  o.n_p = t + g.det_center_channel;
  const float d = vw.angle - theta;
  const float fp = fmaxf(fabsf(cosf(d)) * g.delta_voxel, fabsf(sinf(d)) * g.delta_voxel_row);
  float Wc = mag * (fp / g.delta_det_channel);
  if (!g.curved) Wc = Wc / cosf(theta);
  o.Wc = Wc;
  o.ws = g.area / fp;
  o.mag = mag;
  o.Wr = (mag * g.delta_voxel_slice) / g.delta_det_row;
  o.m0 = (mag * (vw.z_off - g.half_slice_span) + g.det_row_offset) / g.delta_det_row +
         g.det_center_row;
}

// Trapezoid weight min(max((W+1)/2 - |d|, 0), min(1, W)).
__device__ __forceinline__ float trap(float d, float W) {
  return fminf(fmaxf((W + 1.0f) / 2.0f - fabsf(d), 0.0f), fminf(W, 1.0f));
}

// 1/cos(phi) of the ray through a voxel at height z and magnification mag:
// sqrt(1 + (mag*z/sdd)^2), with 1/sdd from the caller and the hardware
// square root (both within about one ulp of the divided, correctly rounded
// form; the forward and back projectors share this function).
__device__ __forceinline__ float inv_cos_phi(float mag, float z, float inv_sdd) {
  const float r = (mag * z) * inv_sdd;
  return __builtin_amdgcn_sqrtf(1.0f + r * r);
}

// Horizontal taps: base = round(n_p) - psf, NH weights (zero off-detector).
template <int NH>
__device__ __forceinline__ void cone_htaps(const ConePixel& q, const mbir_cone_geom& g,
                                           int& base, float (&a)[NH]) {
  base = (int)rintf(q.n_p) - g.psf_radius;
#pragma unroll
  for (int k = 0; k < NH; ++k) {
    const int n = base + k;
    const float A = trap(q.n_p - (float)n, q.Wc) * q.ws;
    a[k] = (n >= 0 && n < g.num_channels) ? A : 0.0f;
  }
}

// Orders a wavefront's LDS writes before its later LDS reads (lanes of one
// wavefront exchanging values; no other wavefront takes part).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

constexpr int kConeRowCap = 256;  // detector rows of one wavefront's shared pass

// ---------------------------------------------------------------------------
// Back projection.  LDS record per (view, pixel): base, NH weights, m0, Wr, mag.
// With one pixel per wavefront (64 slices on the lanes), when weighted or
// with windows wider than 3 x 3, the slices of a (pixel, view) share detector
// rows: the wavefront first forms each row's
// horizontal sum det(m) = sum_t a_t * e(base + t, m) once (lanes = rows,
// coalesced) into LDS, then every slice lane adds its vertical taps
// A_v(m) * det(m) from there -- the same float operations in the same order
// as forming det(m) per slice, so the sums are bit-identical.
template <int RW, int NH, int NV, int CP, bool WEIGHTED, typename TS>
__global__ __launch_bounds__(kCBlock) void cone_back_kernel(
    const TS* __restrict__ sino, const TS* __restrict__ wts,
    const int64_t* __restrict__ pix, int n_pix, const ConeView* __restrict__ views,
    mbir_cone_geom g, int views_per_split, float* __restrict__ out) {
  constexpr int SLOTS = MBIR_WAVE / RW;
  constexpr int ACC = RW >= 4 ? 4 : 2;
  constexpr int PIX_WAVE = SLOTS * ACC;
  constexpr int PIX_WG = (kCBlock / MBIR_WAVE) * PIX_WAVE;
  constexpr int VB = PIX_WG >= 512 ? 1 : 512 / PIX_WG;
  constexpr int REC = NH + 4;
  // The shared-row pass pays where a slice's own loop would load more than
  // nine values per (pixel, view): weighted, or windows wider than 3 x 3
  // (measured: 3 x 3 unweighted is faster per slice).
  constexpr bool ROWS = SLOTS == 1 && (WEIGHTED || NH * NV > 9);
  __shared__ float s_rec[REC * VB * PIX_WG];
  __shared__ float s_zoff[VB];
  __shared__ float s_det[ROWS ? kCBlock / MBIR_WAVE : 1][ROWS ? kConeRowCap : 1];

  const int lane = threadIdx.x & (MBIR_WAVE - 1), wave = threadIdx.x / MBIR_WAVE;
  const int slot = lane / RW, kl = lane % RW;
  const int S = g.num_slices, C = g.num_channels, R = g.num_det_rows;
  // Split view ranges: XCD-aware order, one XCD runs consecutive pixel
  // groups of one split (its sinogram lines stay in one L2).
  int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
  if (gridDim.y > 1) {
    const int nx = gridDim.x, ny = gridDim.y;
    const int l = xcd_remap(blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z), nx * ny * gridDim.z);
    bx = l % nx;
    by = (l / nx) % ny;
    bz = l / (nx * ny);
  }
  const int k = bz * RW + kl;
  const bool k_ok = k < S;
  const float kf = (float)k;
  const float zrel = g.delta_voxel_slice * (kf - g.slice_center);
  const float inv_sdd = 1.0f / g.source_detector_dist;
  const int p_base = bx * PIX_WG;
  const int v_begin = by * views_per_split;
  const int v_end = min(v_begin + views_per_split, g.num_views);
  constexpr int STRIDE = VB * PIX_WG;

  float acc[ACC];
#pragma unroll
  for (int i = 0; i < ACC; ++i) acc[i] = 0.0f;

  for (int vb0 = v_begin; vb0 < v_end; vb0 += VB) {
    const int nvb = min(VB, v_end - vb0);
    __syncthreads();
    if (threadIdx.x < VB) s_zoff[threadIdx.x] = threadIdx.x < nvb ? views[vb0 + threadIdx.x].z_off : 0.0f;
    for (int qd = threadIdx.x; qd < STRIDE; qd += kCBlock) {
      const int j = qd % PIX_WG, b = qd / PIX_WG;
      const int pg = p_base + j;
      int base = 0;
      float a[NH];
      ConePixel q = {0, 1, 0, 1, -1e9f, 1};
      if (b < nvb && pg < n_pix) {
        cone_pixel(pix[pg], views[vb0 + b], g, q);
        cone_htaps<NH>(q, g, base, a);
      } else {
#pragma unroll
        for (int t = 0; t < NH; ++t) a[t] = 0.0f;
      }
      s_rec[qd] = __int_as_float(base);
#pragma unroll
      for (int t = 0; t < NH; ++t) s_rec[(1 + t) * STRIDE + qd] = CP == 2 ? a[t] * a[t] : a[t];
      s_rec[(1 + NH) * STRIDE + qd] = q.m0;
      s_rec[(2 + NH) * STRIDE + qd] = q.Wr;
      s_rec[(3 + NH) * STRIDE + qd] = q.mag;
    }
    __syncthreads();
    // Lane of the last slice below S (wave-uniform; the shared-row pass).
    const int last = min(RW, S - bz * RW) - 1;
    if (ROWS || k_ok) {
      for (int b = 0; b < nvb; ++b) {
        const float z = zrel + s_zoff[b];
        const size_t vofs = (size_t)(vb0 + b) * C * R;
#pragma unroll
        for (int i = 0; i < ACC; ++i) {
          const int qd = b * PIX_WG + wave * PIX_WAVE + i * SLOTS + slot;
          const int base = __float_as_int(s_rec[qd]);
          const float m0 = s_rec[(1 + NH) * STRIDE + qd];
          const float Wr = s_rec[(2 + NH) * STRIDE + qd];
          const float mag = s_rec[(3 + NH) * STRIDE + qd];
          const float ic = inv_cos_phi(mag, z, inv_sdd);
          const float m_p = m0 + Wr * kf;
          const int mc = (int)rintf(m_p) - g.psf_radius;
          float sum = 0.0f;
          bool shared_rows = false;
          if constexpr (ROWS) {
            // mc grows with the slice (Wr > 0): rows lo .. hi cover every
            // lane's taps.  Other spans take the per-slice loop below.
            const int lo = __shfl(mc, 0, MBIR_WAVE);
            const int cnt = __shfl(mc, last, MBIR_WAVE) + NV - lo;
            if (cnt >= NV && cnt <= kConeRowCap) {
              shared_rows = true;
              wave_lds_sync();  // the previous pixel's reads are done
              for (int j = lane; j < cnt; j += MBIR_WAVE) {
                const int mcl = min(max(lo + j, 0), R - 1);
                float det = 0.0f;
#pragma unroll
                for (int th = 0; th < NH; ++th) {
                  const float ah = s_rec[(1 + th) * STRIDE + qd];
                  const int n = min(max(base + th, 0), C - 1);
                  float e = ld_s(sino + vofs + (size_t)n * R + mcl);
                  if (WEIGHTED) e = e * ld_s(wts + vofs + (size_t)n * R + mcl);
                  det = fmaf(ah, e, det);
                }
                s_det[wave][j] = det;
              }
              wave_lds_sync();
              if (k_ok) {
#pragma unroll
                for (int tv = 0; tv < NV; ++tv) {
                  const int mm = mc + tv;
                  float Av = fminf(fmaxf((Wr + 1.0f) / 2.0f - fabsf(m_p - (float)mm), 0.0f),
                                   fminf(Wr, 1.0f)) * ic;
                  Av = (mm >= 0 && mm < R) ? Av : 0.0f;
                  if (CP == 2) Av = Av * Av;
                  const int j = min(max(mm - lo, 0), cnt - 1);
                  sum = fmaf(Av, s_det[wave][j], sum);
                }
              }
            }
          }
          if (!shared_rows && k_ok) {
#pragma unroll
          for (int tv = 0; tv < NV; ++tv) {
            const int mm = mc + tv;
            float Av = fminf(fmaxf((Wr + 1.0f) / 2.0f - fabsf(m_p - (float)mm), 0.0f),
                             fminf(Wr, 1.0f)) * ic;
            Av = (mm >= 0 && mm < R) ? Av : 0.0f;
            if (CP == 2) Av = Av * Av;
            const int mcl = min(max(mm, 0), R - 1);
            float det = 0.0f;
#pragma unroll
            for (int th = 0; th < NH; ++th) {
              const float ah = s_rec[(1 + th) * STRIDE + qd];
              const int n = min(max(base + th, 0), C - 1);
              float e = ld_s(sino + vofs + (size_t)n * R + mcl);
              if (WEIGHTED) e = e * ld_s(wts + vofs + (size_t)n * R + mcl);
              det = fmaf(ah, e, det);
            }
            sum = fmaf(Av, det, sum);
          }
          }
          acc[i] += sum;
        }
      }
    }
  }
  if (k_ok) {
#pragma unroll
    for (int i = 0; i < ACC; ++i) {
      const int pg = p_base + wave * PIX_WAVE + i * SLOTS + slot;
      if (pg < n_pix) out[((size_t)by * n_pix + pg) * S + k] = acc[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Forward projection: binned gather (atomic-free, deterministic).
// Workgroup = (view, kConeCB channels, row chunk of LPC detector rows); lane
// rl of a unit owns detector row m = r0 + rl.  Chunks of the pixel list are
// counting-sorted by channel in LDS with a placement fixed by the list
// (short bins: scatter then insertion sort by position; long bins: wave
// ballot ranks), the pixel's vertical map (m0, Wr, magnification) kept per
// list position.  Unit u gathers the bins of channels c0 + u + k*UNITS in
// rounds: for each entry the lane forms the reference's detector-side
// vertical fan sum for its row -- slices round(k_m) +- bp_psf with
// k_m = (m - m0)/Wr, trapezoid weights, values scaled by 1/cos(phi) -- and
// adds the entry's horizontal weight times it (cone_beam.py:161-234).
// Epilogues as in the parallel projector: write, FP64 line search, or error
// update.
enum ConeFwdMode { CF_WRITE_VCR = 0, CF_WRITE_VRC = 1, CF_LINESEARCH = 2, CF_UPDATE = 3 };

constexpr int kConeCB = 64;     // channels per workgroup
constexpr int kConeShortBin = 24;

template <int NH>
struct ConeBinCfg {
  static constexpr int CH = NH <= 3 ? 1024 : 512;  // pixels per chunk
  static constexpr int PPT = CH / kCBlock;         // pixels per thread
  static constexpr int ENT = CH * NH;              // entry capacity
};

template <int LPC, int NH, int NK, int MODE, bool WEIGHTED, typename TS>
__global__ __launch_bounds__(kCBlock) void cone_fwd_bin_kernel(
    const float* __restrict__ values, const int64_t* __restrict__ pix, int n_pix,
    const ConeView* __restrict__ views, mbir_cone_geom g, TS* __restrict__ sino,
    const TS* __restrict__ wts, const float* __restrict__ alpha,
    double* __restrict__ partials, float* __restrict__ adelta) {
  using Cfg = ConeBinCfg<NH>;
  constexpr int SPW = MBIR_WAVE / LPC;
  constexpr int UNITS = kCBlock / LPC;
  constexpr int CPU = kConeCB / UNITS;  // channels per unit (LPC >= 4)
  constexpr int NW = kCBlock / MBIR_WAVE;
  __shared__ int s_cnt[kConeCB + 1];
  __shared__ int s_off[kConeCB + 1];
  __shared__ int s_wcnt[NW][kConeCB];
  __shared__ int s_ej[Cfg::ENT];
  __shared__ float s_ea[Cfg::ENT];
  __shared__ float s_m0[Cfg::CH], s_wr[Cfg::CH], s_mag[Cfg::CH];

  const int tid = threadIdx.x;
  const int lane = tid & (MBIR_WAVE - 1), wave = tid / MBIR_WAVE;
  const int unit = wave * SPW + lane / LPC, rl = lane % LPC;
  const int v = blockIdx.x, c0 = blockIdx.y * kConeCB, r0 = blockIdx.z * LPC;
  const int S = g.num_slices, C = g.num_channels, R = g.num_det_rows;
  const int m = r0 + rl;
  const bool row_ok = m < R;
  const float mf = (float)m;
  const ConeView vw = views[v];
  const float inv_sdd = 1.0f / g.source_detector_dist;
  if (tid < kConeCB) {
#pragma unroll
    for (int w = 0; w < NW; ++w) s_wcnt[w][tid] = 0;
  }

  float acc[CPU];
#pragma unroll
  for (int k = 0; k < CPU; ++k) acc[k] = 0.0f;

  for (int pc = 0; pc < n_pix; pc += Cfg::CH) {
    // Geometry of pixels pc + tid + q*kCBlock; their horizontal taps landing
    // in the block count.
    int bin[Cfg::PPT][NH];
    float a[Cfg::PPT][NH];
    if (tid < kConeCB) s_cnt[tid] = 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Cfg::PPT; ++q) {
      const int jl = tid + q * kCBlock;
      int base = 0;
      ConePixel cp = {0, 1, 0, 1, -1e9f, 1};
      if (pc + jl < n_pix) {
        cone_pixel(pix[pc + jl], vw, g, cp);
        cone_htaps<NH>(cp, g, base, a[q]);
      } else {
#pragma unroll
        for (int t = 0; t < NH; ++t) a[q][t] = 0.0f;
      }
      s_m0[jl] = cp.m0;
      s_wr[jl] = cp.Wr;
      s_mag[jl] = cp.mag;
#pragma unroll
      for (int t = 0; t < NH; ++t) {
        const int b = base + t - c0;
        const bool in = a[q][t] != 0.0f && b >= 0 && b < kConeCB;
        bin[q][t] = in ? b : -1;
        if (in) atomicAdd(&s_cnt[b], 1);
      }
    }
    __syncthreads();
    int mine = 0;
    if (wave == 0) {  // exclusive scan of the kConeCB == 64 bin counts
      mine = s_cnt[lane];
      int x = mine;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
      }
      s_off[lane] = x - mine;
      if (lane == 63) s_off[kConeCB] = x;
    }
    __syncthreads();
    if (tid < kConeCB) s_cnt[tid] = s_off[tid];  // placement cursor
    if (!__syncthreads_or(tid < kConeCB && mine > kConeShortBin)) {
#pragma unroll
      for (int q = 0; q < Cfg::PPT; ++q) {
#pragma unroll
        for (int t = 0; t < NH; ++t) {
          const int b = bin[q][t];
          if (b >= 0) {
            const int slot = atomicAdd(&s_cnt[b], 1);
            s_ej[slot] = tid + q * kCBlock;
            s_ea[slot] = a[q][t];
          }
        }
      }
      __syncthreads();
      if (tid < kConeCB) {
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
      }
    } else {
#pragma unroll
      for (int q = 0; q < Cfg::PPT; ++q) {
#pragma unroll
        for (int t = 0; t < NH; ++t) {
          const int b = bin[q][t];
          int rank, gsize;
          bool leader;
          wave_bin_rank(b, rank, gsize, leader);
          if (leader) s_wcnt[wave][b] = gsize;
          if (!__syncthreads_or(b >= 0)) continue;
          if (b >= 0) {
            int slot = s_cnt[b] + rank;
            for (int w = 0; w < wave; ++w) slot += s_wcnt[w][b];
            s_ej[slot] = tid + q * kCBlock;
            s_ea[slot] = a[q][t];
          }
          __syncthreads();
          if (tid < kConeCB) {
            int add = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
              add += s_wcnt[w][tid];
              s_wcnt[w][tid] = 0;
            }
            s_cnt[tid] += add;
          }
          __syncthreads();
        }
      }
    }
    __syncthreads();
    if (row_ok) {
      int span[CPU];  // (bin start << 16) | bin length; ENT < 65536
      int rounds = 0;
#pragma unroll
      for (int k = 0; k < CPU; ++k) {
        const int ch = unit + k * UNITS;
        const int e0 = s_off[ch], len = s_off[ch + 1] - e0;
        span[k] = (e0 << 16) | len;
        rounds = max(rounds, len);
      }
      for (int r = 0; r < rounds; ++r) {
#pragma unroll
        for (int k = 0; k < CPU; ++k) {
          if (r < (span[k] & 0xffff)) {
            const int e = (span[k] >> 16) + r;
            const int j = s_ej[e];
            const float m0 = s_m0[j], Wr = s_wr[j], mag = s_mag[j];
            const float* xp = values + (size_t)(pc + j) * S;
            // Detector row -> fractional slice (inverse of the affine map).
            const float k_m = (mf - m0) / Wr;
            const float kc = rintf(k_m);
            const float m_p = Wr * (kc - k_m);
            const float half = (Wr + 1.0f) / 2.0f, lmax = fminf(Wr, 1.0f);
            float det = 0.0f;
#pragma unroll
            for (int tk = 0; tk < NK; ++tk) {
              const int off = tk - (NK - 1) / 2;
              const int ki = (int)kc + off;
              float A = fminf(fmaxf(half - fabsf(m_p + Wr * (float)off), 0.0f), lmax);
              A = (ki >= 0 && ki < S) ? A : 0.0f;
              const int kcl = min(max(ki, 0), S - 1);
              const float z = g.delta_voxel_slice * ((float)kcl - g.slice_center) + vw.z_off;
              const float gv = xp[kcl] * inv_cos_phi(mag, z, inv_sdd);
              det = det + A * gv;
            }
            acc[k] = acc[k] + s_ea[e] * det;
          }
        }
      }
    }
    __syncthreads();
  }

  double lin = 0.0, quad = 0.0;
  const float al = MODE == CF_UPDATE ? alpha[0] : 0.0f;
  if (row_ok) {
#pragma unroll
    for (int k = 0; k < CPU; ++k) {
      const int n = c0 + unit + k * UNITS;
      if (n >= C) continue;
      const float t = acc[k];
      if (MODE == CF_WRITE_VRC) {
        if constexpr (sizeof(TS) == sizeof(float)) sino[((size_t)v * R + m) * C + n] = t;
        continue;
      }
      const size_t o = ((size_t)v * C + n) * R + m;
      if (MODE == CF_WRITE_VCR) {
        if constexpr (sizeof(TS) == sizeof(float)) sino[o] = t;
      } else if (MODE == CF_LINESEARCH) {
        if (adelta) adelta[o] = t;  // A(delta) kept for the error update
        // No error sinogram: the linear term comes from the back projection.
        const float e = sino ? ld_s(sino + o) : 0.0f;
        if (WEIGHTED) {
          const float w = ld_s(wts + o);
          if (sino) lin += (double)((w * e) * t);
          quad += (double)((t * t) * w);
        } else {
          if (sino) lin += (double)(e * t);
          quad += (double)(t * t);
        }
      } else {
        st_s(sino + o, ld_s(sino + o) - al * t);
      }
    }
  }
  if (MODE == CF_LINESEARCH) {
    block_sum2_f64(lin, quad);
    if (tid == 0) {
      const size_t bi = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
      partials[2 * bi] = lin;
      partials[2 * bi + 1] = quad;
    }
  }
}

// ---------------------------------------------------------------------------
static inline int cnext_pow2(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

static int check_cone(const mbir_cone_geom* g) {
  if (!g) return MBIR_ERR_ARG;
  if (g->num_rows <= 0 || g->num_cols <= 0 || g->num_slices <= 0 || g->num_det_rows <= 0 ||
      g->num_channels <= 0 || g->num_views <= 0)
    return MBIR_ERR_SHAPE;
  if (g->psf_radius < 1 || 2 * g->psf_radius + 1 > kCMaxTaps) return MBIR_ERR_GEOMETRY;
  if (g->bp_psf_radius < 1 || 2 * g->bp_psf_radius + 1 > kCMaxTaps) return MBIR_ERR_GEOMETRY;
  return MBIR_OK;
}

static inline int cone_rw(int S) { return S >= 64 ? 64 : cnext_pow2(S); }
static inline int cone_pix_per_wg(int RW) {
  return (kCBlock / MBIR_WAVE) * (MBIR_WAVE / RW) * (RW >= 4 ? 4 : 2);
}

// Detector rows per forward workgroup (one lane each): the next power of
// two >= R, between 4 and 64.
static inline int cone_fwd_lpc(int R) {
  const int p = cnext_pow2(R);
  return p < 4 ? 4 : (p > 64 ? 64 : p);
}

// Sinogram storage: float32, or bfloat16 bits (back with cp == 1, forward
// line search / update).  Tap-window templates: 3, 5 or 9 taps, the
// narrowest that holds the reference's 2*psf+1 (the trapezoid weights vanish
// outside it, so the extra taps of a wider window add exact zeros).
template <int RW, int NH, int NV>
static int cone_back_launch3(const void* sino, const void* w, const int64_t* pix, int P,
                             const ConeView* views, const mbir_cone_geom& g, int cp, int ns,
                             float* out, hipStream_t s, bool bf16) {
  const int pwg = cone_pix_per_wg(RW);
  const int vps = (g.num_views + ns - 1) / ns;
  dim3 grid((P + pwg - 1) / pwg, ns, (g.num_slices + RW - 1) / RW);
  const float* sf = static_cast<const float*>(sino);
  const float* wf = static_cast<const float*>(w);
  const bf16_bits* sh = static_cast<const bf16_bits*>(sino);
  const bf16_bits* wh = static_cast<const bf16_bits*>(w);
  if (bf16) {
    if (w) cone_back_kernel<RW, NH, NV, 1, true, bf16_bits><<<grid, kCBlock, 0, s>>>(sh, wh, pix, P, views, g, vps, out);
    else cone_back_kernel<RW, NH, NV, 1, false, bf16_bits><<<grid, kCBlock, 0, s>>>(sh, wh, pix, P, views, g, vps, out);
  } else if (cp == 1) {
    if (w) cone_back_kernel<RW, NH, NV, 1, true, float><<<grid, kCBlock, 0, s>>>(sf, wf, pix, P, views, g, vps, out);
    else cone_back_kernel<RW, NH, NV, 1, false, float><<<grid, kCBlock, 0, s>>>(sf, wf, pix, P, views, g, vps, out);
  } else {
    if (w) cone_back_kernel<RW, NH, NV, 2, true, float><<<grid, kCBlock, 0, s>>>(sf, wf, pix, P, views, g, vps, out);
    else cone_back_kernel<RW, NH, NV, 2, false, float><<<grid, kCBlock, 0, s>>>(sf, wf, pix, P, views, g, vps, out);
  }
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

template <int NT>
static int cone_back_taps(const void* sino, const void* w, const int64_t* pix, int P,
                          const ConeView* views, const mbir_cone_geom& g, int cp, int ns,
                          float* out, hipStream_t s, bool bf16) {
  switch (cone_rw(g.num_slices)) {
    case 64: return cone_back_launch3<64, NT, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
    case 32: return cone_back_launch3<32, NT, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
    case 16: return cone_back_launch3<16, NT, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
    case 8: return cone_back_launch3<8, NT, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
    case 4: return cone_back_launch3<4, NT, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
    case 2: return cone_back_launch3<2, NT, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
    default: return cone_back_launch3<1, NT, NT>(sino, w, pix, P, views, g, cp, ns, out, s, bf16);
  }
}

struct ConeFwdArgs {
  int mode;
  const float* values;
  const int64_t* pix;
  int P;
  const ConeView* views;
  mbir_cone_geom g;
  void* sino;
  const void* w;
  const float* alpha;
  double* partials;
  hipStream_t s;
  bool bf16;
  float* adelta;  // line search: also store A(delta) (channel-major), or NULL
};

template <int LPC, int NH, int NK, int MODE, bool WEIGHTED, typename TS>
static void cone_fwd_kernel_launch(const ConeFwdArgs& a) {
  dim3 grid(a.g.num_views, (a.g.num_channels + kConeCB - 1) / kConeCB,
            (a.g.num_det_rows + LPC - 1) / LPC);
  cone_fwd_bin_kernel<LPC, NH, NK, MODE, WEIGHTED, TS><<<grid, kCBlock, 0, a.s>>>(
      a.values, a.pix, a.P, a.views, a.g, static_cast<TS*>(a.sino),
      static_cast<const TS*>(a.w), a.alpha, a.partials, a.adelta);
}

template <int LPC, int NH, int NK>
static int cone_fwd_launch3(const ConeFwdArgs& a) {
  if (a.bf16) {
    if (a.mode == CF_LINESEARCH) {
      if (a.w) cone_fwd_kernel_launch<LPC, NH, NK, CF_LINESEARCH, true, bf16_bits>(a);
      else cone_fwd_kernel_launch<LPC, NH, NK, CF_LINESEARCH, false, bf16_bits>(a);
    } else {
      cone_fwd_kernel_launch<LPC, NH, NK, CF_UPDATE, false, bf16_bits>(a);
    }
  } else {
    switch (a.mode) {
      case CF_WRITE_VCR: cone_fwd_kernel_launch<LPC, NH, NK, CF_WRITE_VCR, false, float>(a); break;
      case CF_WRITE_VRC: cone_fwd_kernel_launch<LPC, NH, NK, CF_WRITE_VRC, false, float>(a); break;
      case CF_LINESEARCH:
        if (a.w) cone_fwd_kernel_launch<LPC, NH, NK, CF_LINESEARCH, true, float>(a);
        else cone_fwd_kernel_launch<LPC, NH, NK, CF_LINESEARCH, false, float>(a);
        break;
      default: cone_fwd_kernel_launch<LPC, NH, NK, CF_UPDATE, false, float>(a);
    }
  }
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

template <int NH, int NK>
static int cone_fwd_rc_sw(const ConeFwdArgs& a) {
  switch (cone_fwd_lpc(a.g.num_det_rows)) {
    case 64: return cone_fwd_launch3<64, NH, NK>(a);
    case 32: return cone_fwd_launch3<32, NH, NK>(a);
    case 16: return cone_fwd_launch3<16, NH, NK>(a);
    case 8: return cone_fwd_launch3<8, NH, NK>(a);
    default: return cone_fwd_launch3<4, NH, NK>(a);
  }
}

static int cone_fwd(const ConeFwdArgs& a) {
  if (a.bf16 && (a.mode == CF_WRITE_VCR || a.mode == CF_WRITE_VRC)) return MBIR_ERR_ARG;
  const int nh = 2 * a.g.psf_radius + 1;
  const bool nk3 = 2 * a.g.bp_psf_radius + 1 <= 3;
  if (nh <= 3) return nk3 ? cone_fwd_rc_sw<3, 3>(a) : cone_fwd_rc_sw<3, 9>(a);
  if (nh <= 5) return nk3 ? cone_fwd_rc_sw<5, 3>(a) : cone_fwd_rc_sw<5, 9>(a);
  return nk3 ? cone_fwd_rc_sw<9, 3>(a) : cone_fwd_rc_sw<9, 9>(a);
}

static int cone_back_any(const void* sino, const void* w, const int64_t* pix, int32_t P,
                         const float* view_table, const mbir_cone_geom* g, int32_t cp, int32_t ns,
                         float* out, mbir_stream_t stream, bool bf16);

}  // namespace mbir

using namespace mbir;

int mbir::cone_back_any(const void* sino, const void* w, const int64_t* pix, int32_t P,
                          const float* view_table, const mbir_cone_geom* g, int32_t cp,
                          int32_t ns, float* out, mbir_stream_t stream, bool bf16) {
  int st = check_cone(g);
  if (st) return st;
  if (!sino || !pix || !view_table || !out || P < 0) return MBIR_ERR_ARG;
  if ((cp != 1 && cp != 2) || ns < 1 || ns > g->num_views) return MBIR_ERR_ARG;
  if (P == 0) return MBIR_OK;
  const ConeView* views = reinterpret_cast<const ConeView*>(view_table);
  hipStream_t s = (hipStream_t)stream;
  const int nt = 2 * g->psf_radius + 1;
  if (nt <= 3) return cone_back_taps<3>(sino, w, pix, P, views, *g, cp, ns, out, s, bf16);
  if (nt <= 5) return cone_back_taps<5>(sino, w, pix, P, views, *g, cp, ns, out, s, bf16);
  return cone_back_taps<9>(sino, w, pix, P, views, *g, cp, ns, out, s, bf16);
}

extern "C" {

int32_t mbir_cone_back_splits(const mbir_cone_geom* g, int32_t P) {
  if (!g || P <= 0) return 1;
  const int RW = cone_rw(g->num_slices);
  const int pwg = cone_pix_per_wg(RW);
  const long wg = (long)((P + pwg - 1) / pwg) * ((g->num_slices + RW - 1) / RW);
  long ns = (2048 + wg - 1) / wg;
  long max_ns = g->num_views / 8;
  if (ns > max_ns) ns = max_ns;
  if (ns < 1) ns = 1;
  return (int32_t)ns;
}

int mbir_cone_back(const float* sino, const float* w, const int64_t* pix, int32_t P,
                   const float* view_table, const mbir_cone_geom* g, int32_t cp, int32_t ns,
                   float* out, mbir_stream_t stream) {
  return cone_back_any(sino, w, pix, P, view_table, g, cp, ns, out, stream, false);
}

int mbir_cone_back_bf16(const uint16_t* sino, const uint16_t* w, const int64_t* pix, int32_t P,
                        const float* view_table, const mbir_cone_geom* g, int32_t ns, float* out,
                        mbir_stream_t stream) {
  return cone_back_any(sino, w, pix, P, view_table, g, 1, ns, out, stream, true);
}

int mbir_cone_forward(const float* values, const int64_t* pix, int32_t P,
                      const float* view_table, const mbir_cone_geom* g, int32_t layout,
                      float* out, mbir_stream_t stream) {
  int st = check_cone(g);
  if (st) return st;
  if (!values || !pix || !view_table || !out || P < 0) return MBIR_ERR_ARG;
  if (layout != MBIR_LAYOUT_VRC && layout != MBIR_LAYOUT_VCR) return MBIR_ERR_ARG;
  ConeFwdArgs a{layout == MBIR_LAYOUT_VRC ? CF_WRITE_VRC : CF_WRITE_VCR, values, pix, P,
                reinterpret_cast<const ConeView*>(view_table), *g, out, nullptr, nullptr,
                nullptr, (hipStream_t)stream, false, nullptr};
  return cone_fwd(a);
}

int32_t mbir_cone_fwd_partials_count(const mbir_cone_geom* g) {
  if (!g || check_cone(g)) return 0;
  const int lpc = cone_fwd_lpc(g->num_det_rows);
  return g->num_views * ((g->num_channels + kConeCB - 1) / kConeCB) *
         ((g->num_det_rows + lpc - 1) / lpc);
}

static int cone_ls_any(const float* delta, const int64_t* pix, int32_t P, const float* view_table,
                       const mbir_cone_geom* g, const void* error_vcr, const void* w,
                       float* adelta, double* partials, mbir_stream_t stream, bool bf16) {
  int st = check_cone(g);
  if (st) return st;
  if (!delta || !pix || !view_table || !partials || P < 0) return MBIR_ERR_ARG;
  ConeFwdArgs a{CF_LINESEARCH, delta, pix, P, reinterpret_cast<const ConeView*>(view_table), *g,
                const_cast<void*>(error_vcr), w, nullptr, partials, (hipStream_t)stream, bf16,
                adelta};
  return cone_fwd(a);
}

static int cone_update_any(const float* delta, const int64_t* pix, int32_t P,
                           const float* view_table, const mbir_cone_geom* g, const float* alpha,
                           void* error_vcr, mbir_stream_t stream, bool bf16) {
  int st = check_cone(g);
  if (st) return st;
  if (!delta || !pix || !view_table || !error_vcr || !alpha || P < 0) return MBIR_ERR_ARG;
  ConeFwdArgs a{CF_UPDATE, delta, pix, P, reinterpret_cast<const ConeView*>(view_table), *g,
                error_vcr, nullptr, alpha, nullptr, (hipStream_t)stream, bf16, nullptr};
  return cone_fwd(a);
}

int mbir_cone_fwd_linesearch(const float* delta, const int64_t* pix, int32_t P,
                             const float* view_table, const mbir_cone_geom* g,
                             const float* error_vcr, const float* w, float* adelta_vcr,
                             double* partials, mbir_stream_t stream) {
  return cone_ls_any(delta, pix, P, view_table, g, error_vcr, w, adelta_vcr, partials, stream,
                     false);
}

int mbir_cone_fwd_linesearch_bf16(const float* delta, const int64_t* pix, int32_t P,
                                  const float* view_table, const mbir_cone_geom* g,
                                  const uint16_t* error_vcr, const uint16_t* w,
                                  float* adelta_vcr, double* partials, mbir_stream_t stream) {
  return cone_ls_any(delta, pix, P, view_table, g, error_vcr, w, adelta_vcr, partials, stream,
                     true);
}

int mbir_cone_fwd_update(const float* delta, const int64_t* pix, int32_t P,
                         const float* view_table, const mbir_cone_geom* g, const float* alpha,
                         float* error_vcr, mbir_stream_t stream) {
  return cone_update_any(delta, pix, P, view_table, g, alpha, error_vcr, stream, false);
}

int mbir_cone_fwd_update_bf16(const float* delta, const int64_t* pix, int32_t P,
                              const float* view_table, const mbir_cone_geom* g,
                              const float* alpha, uint16_t* error_vcr, mbir_stream_t stream) {
  return cone_update_any(delta, pix, P, view_table, g, alpha, error_vcr, stream, true);
}

}  // extern "C"
