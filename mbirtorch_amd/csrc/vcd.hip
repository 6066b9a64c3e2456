// VCD subset-step kernels for MI355X (gfx950): prior + direction, the
// on-device step size, positivity, the recon update, and FP64 reductions.
//
// Elementwise float32 arithmetic follows the reference's order
// (qggmrf.py:35-183, tomography_model.py:54-87, 2508-2657); every sum that
// decides the step size or a reported statistic is accumulated in FP64 with
// a fixed reduction tree, so a subset step is deterministic and needs no
// host synchronisation (alpha stays on the device).
#include "common.h"

namespace mbir {

constexpr int kThreads = 256;
constexpr float kEps32 = 1.1920928955078125e-07f;

// a^b for a > 0 from the hardware base-2 logarithm and exponential (each
// within about an ulp; the product's relative error is ~|b log2 a| * 2^-23).
__device__ __forceinline__ float pow_pos(float a, float b) {
  return __builtin_amdgcn_exp2f(b * __builtin_amdgcn_logf(a));
}

// 2*b_tilde(delta) without the direction weight: rho'(delta)/delta of the
// qGGMRF potential with |delta| floored at a_min > 0 (qggmrf.py:35-57).
__device__ __forceinline__ float qg_b_tilde(float delta, const mbir_vcd_params& p) {
  const float ad = fmaxf(fabsf(delta), p.a_min);
  const float ds = ad / p.t_sigma;
  const float u = pow_pos(ds, p.q_minus_p);
  const float pm2 = pow_pos(ad, p.p_minus_2);
  float num = pm2 / p.two_sigma_p;
  num = (num * u) * (p.q_over_p + u);
  const float den = 1.0f + u;
  return num / (den * den);
}

// Sum of the view-split back partials of element i, split by split.  With
// G > 1 the G consecutive lanes of a group share element i: lane `sub` sums
// splits sub, sub+G, ... in order and an xor butterfly combines the G sums
// (commutative pairings, so every lane of the group holds the same bits) --
// small subsets then wait for ns/G loads instead of ns.
template <int G>
__device__ __forceinline__ float split_sum(const float* __restrict__ back_parts, int nsplit,
                                           int64_t total, int64_t i, int sub) {
  float bsum = 0.0f;
  int k = sub;
  if (k < nsplit) {
    bsum = back_parts[(size_t)k * total + i];
    k += G;
  }
  // Loads go out in batches of eight so their latencies overlap.
  for (; k + 7 * G < nsplit; k += 8 * G) {
    float t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = back_parts[(size_t)(k + u * G) * total + i];
#pragma unroll
    for (int u = 0; u < 8; ++u) bsum += t[u];
  }
  for (; k < nsplit; k += G) bsum += back_parts[(size_t)k * total + i];
#pragma unroll
  for (int m = 1; m < G; m <<= 1) bsum += __shfl_xor(bsum, m, G);
  return bsum;
}

// Prior gradient/Hessian (qGGMRF surrogate or proximal prior) and forward
// gradient/Hessian at subset element i = (p, s).
template <int G>
__device__ __forceinline__ void vcd_terms(const float* __restrict__ x,
                                          const int64_t* __restrict__ pix, int64_t i, int sub,
                                          int64_t total, const float* __restrict__ back_parts,
                                          int nsplit, const float* __restrict__ hess,
                                          const float* __restrict__ prox,
                                          const mbir_vcd_params& prm, float& fg, float& gp,
                                          float& fh, float& hp, float& bsum_out) {
  const int S = prm.num_slices, Nc = prm.num_cols, Nr = prm.num_rows;
  const int p = (int)(i / S), s = (int)(i - (int64_t)p * S);
  const int64_t pi = pix[p];
  const float xs = x[pi * S + s];
  if (prm.use_prox) {
    gp = prm.inv_sigma_prox2 * (xs - prox[pi * S + s]);
    hp = prm.inv_sigma_prox2;
  } else {
    // Slice neighbours: deltas v[s]-v[s-1] and v[s+1]-v[s], zero at the
    // reflected ends (qggmrf.py:135-156).
    const float dm = s > 0 ? xs - x[pi * S + s - 1] : 0.0f;
    const float dp = s < S - 1 ? x[pi * S + s + 1] - xs : 0.0f;
    const float cm = 2.0f * qg_b_tilde(dm, prm);
    const float cp = 2.0f * qg_b_tilde(dp, prm);
    gp = (-prm.b[4]) * (cp * dp) + prm.b[5] * (cm * dm);
    hp = prm.b[4] * cp + prm.b[5] * cm;
    // In-plane neighbours, indices clamped to the grid (qggmrf.py:158-181).
    const int row = (int)(pi / Nc), col = (int)(pi - (int64_t)row * Nc);
    const int nbr[4] = {min(row + 1, Nr - 1) * Nc + col, max(row - 1, 0) * Nc + col,
                        row * Nc + min(col + 1, Nc - 1), row * Nc + max(col - 1, 0)};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float d = xs - x[(int64_t)nbr[k] * S + s];
      const float c2 = prm.two_b[k] * qg_b_tilde(d, prm);
      gp = gp + c2 * d;
      hp = hp + c2;
    }
  }
  const float bsum = split_sum<G>(back_parts, nsplit, total, i, sub);
  fg = prm.neg_fm_constant * bsum;
  fh = prm.fm_constant * hess[pi * S + s];
  bsum_out = bsum;
}

// G lanes per element (split_sum); lane 0 of each group writes delta and
// accumulates the partials.
template <int G>
__global__ __launch_bounds__(kThreads) void vcd_direction_kernel(
    const float* __restrict__ x, const int64_t* __restrict__ pix, int P,
    const float* __restrict__ back_parts, int nsplit, const float* __restrict__ hess,
    const float* __restrict__ prox, mbir_vcd_params prm, float* __restrict__ delta,
    double* __restrict__ partials) {
  const int64_t total = (int64_t)P * prm.num_slices;
  double pl = 0.0, pq = 0.0, bl = 0.0, unused = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < total * G;
       t += (int64_t)gridDim.x * kThreads) {
    const int64_t i = t / G;
    const int sub = (int)(t % G);
    float fg, gp, fh, hp, bs;
    vcd_terms<G>(x, pix, i, sub, total, back_parts, nsplit, hess, prox, prm, fg, gp, fh, hp, bs);
    const float d = -((fg + gp) / (fh + hp));
    if (sub == 0) {
      delta[i] = d;
      pl += (double)(gp * d);
      pq += (double)((hp * d) * d);
      bl += (double)(bs * d);
    }
  }
  block_sum2_f64(pl, pq);
  block_sum2_f64(bl, unused);
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x] = pl;
    partials[3 * blockIdx.x + 1] = pq;
    partials[3 * blockIdx.x + 2] = bl;
  }
}

// DC damping pass 1: g, h_inv and the prior terms per element.
template <int G>
__global__ __launch_bounds__(kThreads) void vcd_dc_terms_kernel(
    const float* __restrict__ x, const int64_t* __restrict__ pix, int P,
    const float* __restrict__ back_parts, int nsplit, const float* __restrict__ hess,
    const float* __restrict__ prox, mbir_vcd_params prm, float* __restrict__ scratch) {
  const int64_t total = (int64_t)P * prm.num_slices;
  for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < total * G;
       t += (int64_t)gridDim.x * kThreads) {
    const int64_t i = t / G;
    const int sub = (int)(t % G);
    float fg, gp, fh, hp, bs;
    vcd_terms<G>(x, pix, i, sub, total, back_parts, nsplit, hess, prox, prm, fg, gp, fh, hp, bs);
    if (sub != 0) continue;
    scratch[i] = fg + gp;                   // g
    scratch[total + i] = 1.0f / (fh + hp);  // h_inv
    scratch[2 * total + i] = gp;
    scratch[3 * total + i] = hp;
    scratch[4 * total + i] = bs;            // A^T W e
  }
}

// DC damping pass 2: one workgroup per slice, gbar_k in FP64 (fixed order),
// then pass 3 forms delta.
__global__ __launch_bounds__(kThreads) void vcd_dc_gbar_kernel(const float* __restrict__ scratch,
                                                               int P, int S,
                                                               double* __restrict__ gbar) {
  const int s = blockIdx.x;
  const int64_t total = (int64_t)P * S;
  double num = 0.0, den = 0.0;
  for (int p = threadIdx.x; p < P; p += kThreads) {
    const size_t i = (size_t)p * S + s;
    const float h = scratch[total + i];
    num += (double)(h * scratch[i]);
    den += (double)h;
  }
  block_sum2_f64(num, den);
  if (threadIdx.x == 0) gbar[s] = num / den;
}

__global__ __launch_bounds__(kThreads) void vcd_dc_delta_kernel(
    const float* __restrict__ scratch, const double* __restrict__ gbar,
    const float* __restrict__ prof, int P, int S, float* __restrict__ delta,
    double* __restrict__ partials) {
  const int64_t total = (int64_t)P * S;
  double pl = 0.0, pq = 0.0, bl = 0.0, unused = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kThreads) {
    const int s = (int)(i % S);
    const float g = scratch[i], h = scratch[total + i];
    const float gp = scratch[2 * total + i], hp = scratch[3 * total + i];
    const float shift = (1.0f - prof[s]) * (float)gbar[s];
    const float d = -(g - shift) * h;
    delta[i] = d;
    pl += (double)(gp * d);
    pq += (double)((hp * d) * d);
    bl += (double)(scratch[4 * total + i] * d);
  }
  block_sum2_f64(pl, pq);
  block_sum2_f64(bl, unused);
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x] = pl;
    partials[3 * blockIdx.x + 1] = pq;
    partials[3 * blockIdx.x + 2] = bl;
  }
}

// The step size from the prior partials pp (npp pairs) and the line-search
// partials fp (nfp pairs), FP64 in a fixed order; every thread of the
// workgroup returns the same value (tomography_model.py:2620-2632).
__device__ float vcd_alpha_value(const double* __restrict__ pp, int npp,
                                 const double* __restrict__ fp, int nfp, double fm,
                                 double max_alpha) {
  __shared__ float s_alpha;
  double pl = 0.0, pq = 0.0, bl = 0.0, lin = 0.0, quad = 0.0, unused = 0.0;
  for (int i = threadIdx.x; i < npp; i += kThreads) {
    pl += pp[3 * i];
    pq += pp[3 * i + 1];
    bl += pp[3 * i + 2];
  }
  for (int i = threadIdx.x; i < nfp; i += kThreads) { lin += fp[2 * i]; quad += fp[2 * i + 1]; }
  block_sum2_f64(pl, pq);
  block_sum2_f64(lin, quad);
  block_sum2_f64(bl, unused);
  if (threadIdx.x == 0) {
    // <W e, A delta>: from the line search, or (when it ran without the
    // error sinogram) as <A^T W e, delta> from the direction partials.
    lin = lin + bl;
    const double eps = (double)kEps32;
    double a = (fm * lin - pl) / (fm * quad + pq + eps);
    if (!(a >= eps)) a = eps;  // also maps NaN to eps, like a clamp of a NaN-free value
    if (a > max_alpha) a = max_alpha;
    s_alpha = (float)a;
  }
  __syncthreads();
  return s_alpha;
}

__global__ __launch_bounds__(kThreads) void vcd_alpha_kernel(
    const double* __restrict__ pp, int npp, const double* __restrict__ fp, int nfp,
    double fm, double max_alpha, float* __restrict__ alpha, double* __restrict__ alpha_sum) {
  const float af = vcd_alpha_value(pp, npp, fp, nfp, fm, max_alpha);
  if (threadIdx.x == 0) {
    alpha[0] = af;
    alpha_sum[0] += (double)af;
  }
}

__global__ void vcd_positivity_kernel(const float* __restrict__ x, const int64_t* __restrict__ pix,
                                      int P, int S, const float* __restrict__ alpha,
                                      float* __restrict__ delta) {
  const float pc = 1.0f / (alpha[0] + kEps32);
  const int64_t total = (int64_t)P * S;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)(i / S), s = (int)(i - (int64_t)p * S);
    delta[i] = fmaxf(-pc * x[pix[p] * S + s], delta[i]);
  }
}

// x[pix[p]*S + s] += alpha*delta[p, s], fused with the per-workgroup stats
// part[b*(S+1) + s] = sum_p (alpha*delta[p, s])^2 and
// part[b*(S+1) + S] = sum |alpha*delta| over the workgroup's pixel range
// (ACC: added to what part holds instead of stored).
// Lanes map to slices (coalesced cylinders), waves to pixels.
template <bool ACC>
__device__ __forceinline__ void vcd_apply_block(float* __restrict__ x,
                                                const int64_t* __restrict__ pix, int P, int S,
                                                const float* __restrict__ delta, float a, int ppb,
                                                int b, double* __restrict__ part) {
  __shared__ double s_red[kThreads / 64][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int NW = kThreads / 64;
  const int p0 = b * ppb, p1 = min(p0 + ppb, P);
  double* row = part + (size_t)b * (S + 1);
  double ell1 = 0.0;
  for (int sc = 0; sc < S; sc += 64) {
    const int s = sc + lane;
    double sq = 0.0;
    if (s < S) {
      for (int p = p0 + wave; p < p1; p += NW) {
        const float ds = a * delta[(size_t)p * S + s];
        // A subset may list a pixel twice (partition padding); both updates
        // land, as with index_add_.  The duplicates carry identical deltas,
        // so the two atomic adds commute exactly: deterministic.
        atomicAdd(&x[pix[p] * S + s], ds);
        sq += (double)(ds * ds);
        ell1 += (double)fabsf(ds);
      }
    }
    s_red[wave][lane] = sq;
    __syncthreads();
    if (wave == 0 && s < S) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < NW; ++w) t += s_red[w][lane];
      row[s] = ACC ? row[s] + t : t;
    }
    __syncthreads();
  }
  double unused = 0.0;
  block_sum2_f64(ell1, unused);
  if (threadIdx.x == 0) row[S] = ACC ? row[S] + ell1 : ell1;
}

__global__ __launch_bounds__(kThreads) void vcd_apply_kernel(
    float* __restrict__ x, const int64_t* __restrict__ pix, int P, int S,
    const float* __restrict__ delta, const float* __restrict__ alpha, int ppb,
    double* __restrict__ part) {
  vcd_apply_block<false>(x, pix, P, S, delta, alpha[0], ppb, blockIdx.x, part);
}

// acc[c] += sum_b part[b*(S+1) + c], one workgroup per column c, fixed order
// (CLEAR: the column is zeroed after it is read).
template <bool CLEAR>
__global__ __launch_bounds__(kThreads) void vcd_stats_final_kernel(
    double* __restrict__ part, int nb, int S, double* __restrict__ acc) {
  const int c = blockIdx.x;
  double v = 0.0, unused = 0.0;
  for (int b = threadIdx.x; b < nb; b += kThreads) {
    v += part[(size_t)b * (S + 1) + c];
    if (CLEAR) part[(size_t)b * (S + 1) + c] = 0.0;
  }
  block_sum2_f64(v, unused);
  if (threadIdx.x == 0) acc[c] += v;
}

// The end of a subset step in one launch.  With dir_parts set, every
// workgroup first forms the step size from the prior and line-search
// partials (the same value everywhere; workgroup 0 records it in alpha and
// alpha_sum); otherwise alpha[0] holds it.  Workgroups [0, nb_update) then
// stream the error update e -= alpha*A(delta) (the same float32 operations
// as the projector's update epilogue); the others apply x[pix] +=
// alpha*delta and add their statistics into the persistent per-workgroup
// rows of part, which mbir_vcd_block_stats_reduce folds into the
// iteration's statistics.
struct FinishArgs {
  void* e;
  const float* adelta;
  int64_t n_sino;
  int nb_update;
  float* x;
  const int64_t* pix;
  int P, S, ppb;
  const float* delta;
  float* alpha;
  double* alpha_sum;
  const double* dir_parts;
  int n_dir;
  const double* fwd_parts;
  int n_fwd;
  double fm, max_alpha;
  double* part;
};

// Four consecutive sinogram entries (16-byte float / 8-byte bfloat16 words).
__device__ __forceinline__ void ld4(const float* p, float (&v)[4]) {
  const float4 t = *reinterpret_cast<const float4*>(p);
  v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
}
__device__ __forceinline__ void ld4(const bf16_bits* p, float (&v)[4]) {
  const uint2 t = *reinterpret_cast<const uint2*>(p);
  v[0] = __uint_as_float(t.x << 16);
  v[1] = __uint_as_float(t.x & 0xffff0000u);
  v[2] = __uint_as_float(t.y << 16);
  v[3] = __uint_as_float(t.y & 0xffff0000u);
}
__device__ __forceinline__ void st4(float* p, const float (&v)[4]) {
  *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void st4(bf16_bits* p, const float (&v)[4]) {
  uint2 t;
  t.x = (uint32_t)f32_to_bf16(v[0]) | ((uint32_t)f32_to_bf16(v[1]) << 16);
  t.y = (uint32_t)f32_to_bf16(v[2]) | ((uint32_t)f32_to_bf16(v[3]) << 16);
  *reinterpret_cast<uint2*>(p) = t;
}

template <typename TS, bool VEC>
__global__ __launch_bounds__(kThreads) void vcd_finish_kernel(FinishArgs a) {
  float al;
  if (a.dir_parts) {
    al = vcd_alpha_value(a.dir_parts, a.n_dir, a.fwd_parts, a.n_fwd, a.fm, a.max_alpha);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      a.alpha[0] = al;
      a.alpha_sum[0] += (double)al;
    }
  } else {
    al = a.alpha[0];
  }
  if ((int)blockIdx.x < a.nb_update) {
    TS* e = static_cast<TS*>(a.e);
    const int64_t stride = (int64_t)a.nb_update * kThreads;
    if (VEC) {  // n_sino % 4 == 0, aligned rows
      for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < a.n_sino / 4;
           i += stride) {
        float ev[4], dv[4];
        ld4(e + 4 * i, ev);
        ld4(a.adelta + 4 * i, dv);
#pragma unroll
        for (int q = 0; q < 4; ++q) ev[q] = ev[q] - al * dv[q];
        st4(e + 4 * i, ev);
      }
    } else {
      for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < a.n_sino; i += stride)
        st_s(e + i, ld_s(e + i) - al * a.adelta[i]);
    }
    return;  // the whole workgroup leaves: no barrier is skipped
  }
  vcd_apply_block<true>(a.x, a.pix, a.P, a.S, a.delta, al, a.ppb, blockIdx.x - a.nb_update,
                        a.part);
}

static inline int apply_blocks(int P) {
  int nb = (P + 7) / 8;
  return nb > 1024 ? 1024 : (nb < 1 ? 1 : nb);
}

// e -= alpha * A(delta), streaming over the stored projection (the same
// float32 operations as the projector's update epilogue).
template <typename TS>
__global__ __launch_bounds__(kThreads) void error_update_kernel(
    TS* __restrict__ e, const float* __restrict__ adelta, const float* __restrict__ alpha,
    int64_t n) {
  const float al = alpha[0];
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    st_s(e + i, ld_s(e + i) - al * adelta[i]);
}

// Storage conversion (float32 <-> bfloat16, round to nearest even).
template <typename TI, typename TO>
__global__ __launch_bounds__(kThreads) void convert_kernel(const TI* __restrict__ in,
                                                           TO* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    st_s(out + i, ld_s(in + i));
}

// ---- reductions ------------------------------------------------------------
constexpr int kRedItems = 8;  // elements per thread per block pass

template <typename TS>
__global__ __launch_bounds__(kThreads) void sumsq_partial_kernel(const TS* __restrict__ a,
                                                                 const TS* __restrict__ w,
                                                                 int64_t n,
                                                                 double* __restrict__ part) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    const float v = ld_s(a + i);
    const float vv = v * v;
    if (w) {
      const float wv = ld_s(w + i);
      s0 += (double)(vv * wv);
      s2 += (double)wv;
    } else {
      s0 += (double)vv;
      s2 += 1.0;
    }
    s1 += (double)vv;
  }
  double t = s2, unused = 0.0;
  block_sum2_f64(s0, s1);
  block_sum2_f64(t, unused);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s0;
    part[3 * blockIdx.x + 1] = s1;
    part[3 * blockIdx.x + 2] = t;
  }
}

__global__ __launch_bounds__(kThreads) void abs_dot_partial_kernel(const float* __restrict__ a,
                                                                   const float* __restrict__ b,
                                                                   int64_t n,
                                                                   double* __restrict__ part) {
  double s = 0.0, unused = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    s += b ? (double)a[i] * (double)b[i] : (double)fabsf(a[i]);
  }
  block_sum2_f64(s, unused);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s;
    part[3 * blockIdx.x + 1] = 0.0;
    part[3 * blockIdx.x + 2] = 0.0;
  }
}

__global__ __launch_bounds__(kThreads) void final3_kernel(const double* __restrict__ part, int nb,
                                                          int nout, double* __restrict__ out) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, unused = 0.0;
  for (int i = threadIdx.x; i < nb; i += kThreads) {
    s0 += part[3 * i];
    s1 += part[3 * i + 1];
    s2 += part[3 * i + 2];
  }
  block_sum2_f64(s0, s1);
  block_sum2_f64(s2, unused);
  if (threadIdx.x == 0) {
    out[0] = s0;
    if (nout > 1) out[1] = s1;
    if (nout > 2) out[2] = s2;
  }
}

static inline int red_blocks(int64_t n) {
  int64_t b = (n + (int64_t)kThreads * kRedItems - 1) / ((int64_t)kThreads * kRedItems);
  if (b < 1) b = 1;
  if (b > 2048) b = 2048;
  return (int)b;
}

}  // namespace mbir

using namespace mbir;

extern "C" {

// Lanes per element of the direction kernels: small subsets spread each
// element's split sum over a group of lanes.
static inline int dir_group(int32_t P, int32_t S) {
  const int64_t n = (int64_t)P * S;
  return n <= 16384 ? 8 : (n <= 65536 ? 4 : 1);
}

int32_t mbir_vcd_direction_partials_count(int32_t P, int32_t S) {
  // One partial triple per workgroup; large subsets stride over at most 1024.
  const int64_t nb = ((int64_t)P * S * dir_group(P, S) + kThreads - 1) / kThreads;
  return (int32_t)(nb > 1024 ? 1024 : nb);
}

int mbir_vcd_direction(const float* x, const int64_t* pix, int32_t P, const float* back_parts,
                       int32_t nsplit, const float* hess, const float* prox,
                       const mbir_vcd_params* prm, float* delta, double* partials,
                       mbir_stream_t stream) {
  if (!x || !pix || !back_parts || !hess || !prm || !delta || !partials) return MBIR_ERR_ARG;
  if (prm->use_prox && !prox) return MBIR_ERR_ARG;
  if (P <= 0 || nsplit < 1 || prm->num_slices <= 0) return MBIR_ERR_SHAPE;
  const int nb = mbir_vcd_direction_partials_count(P, prm->num_slices);
  hipStream_t s = (hipStream_t)stream;
  switch (dir_group(P, prm->num_slices)) {
    case 8:
      vcd_direction_kernel<8><<<nb, kThreads, 0, s>>>(x, pix, P, back_parts, nsplit, hess, prox,
                                                      *prm, delta, partials);
      break;
    case 4:
      vcd_direction_kernel<4><<<nb, kThreads, 0, s>>>(x, pix, P, back_parts, nsplit, hess, prox,
                                                      *prm, delta, partials);
      break;
    default:
      vcd_direction_kernel<1><<<nb, kThreads, 0, s>>>(x, pix, P, back_parts, nsplit, hess, prox,
                                                      *prm, delta, partials);
  }
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int64_t mbir_vcd_direction_dc_scratch(int32_t P, int32_t S) {
  // 5 float arrays of P*S plus S doubles (as 2*S floats), rounded up.
  return 5 * (int64_t)P * S + 2 * (int64_t)S + 4;
}

int mbir_vcd_direction_dc(const float* x, const int64_t* pix, int32_t P, const float* back_parts,
                          int32_t nsplit, const float* hess, const float* prox,
                          const mbir_vcd_params* prm, const float* prof, float* scratch,
                          float* delta, double* partials, mbir_stream_t stream) {
  if (!x || !pix || !back_parts || !hess || !prm || !prof || !scratch || !delta || !partials)
    return MBIR_ERR_ARG;
  if (prm->use_prox && !prox) return MBIR_ERR_ARG;
  if (P <= 0 || nsplit < 1 || prm->num_slices <= 0) return MBIR_ERR_SHAPE;
  hipStream_t s = (hipStream_t)stream;
  const int S = prm->num_slices;
  const int nb = mbir_vcd_direction_partials_count(P, S);
  const int64_t total = (int64_t)P * S;
  // gbar lives after the five float arrays, 8-byte aligned.
  uintptr_t gaddr = reinterpret_cast<uintptr_t>(scratch + 5 * total);
  gaddr = (gaddr + 7) & ~(uintptr_t)7;
  double* gbar = reinterpret_cast<double*>(gaddr);
  switch (dir_group(P, S)) {
    case 8:
      vcd_dc_terms_kernel<8><<<nb, kThreads, 0, s>>>(x, pix, P, back_parts, nsplit, hess, prox,
                                                     *prm, scratch);
      break;
    case 4:
      vcd_dc_terms_kernel<4><<<nb, kThreads, 0, s>>>(x, pix, P, back_parts, nsplit, hess, prox,
                                                     *prm, scratch);
      break;
    default:
      vcd_dc_terms_kernel<1><<<nb, kThreads, 0, s>>>(x, pix, P, back_parts, nsplit, hess, prox,
                                                     *prm, scratch);
  }
  MBIR_CHECK_LAUNCH();
  vcd_dc_gbar_kernel<<<S, kThreads, 0, s>>>(scratch, P, S, gbar);
  MBIR_CHECK_LAUNCH();
  vcd_dc_delta_kernel<<<nb, kThreads, 0, s>>>(scratch, gbar, prof, P, S, delta, partials);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int mbir_vcd_alpha(const double* pp, int32_t npp, const double* fp, int32_t nfp,
                   const mbir_vcd_params* prm, float* alpha, double* alpha_sum,
                   mbir_stream_t stream) {
  if (!pp || !fp || !prm || !alpha || !alpha_sum || npp < 0 || nfp < 0) return MBIR_ERR_ARG;
  vcd_alpha_kernel<<<1, kThreads, 0, (hipStream_t)stream>>>(
      pp, npp, fp, nfp, (double)prm->fm_constant, (double)prm->max_alpha, alpha, alpha_sum);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int mbir_vcd_positivity(const float* x, const int64_t* pix, int32_t P, int32_t S,
                        const float* alpha, float* delta, mbir_stream_t stream) {
  if (!x || !pix || !alpha || !delta) return MBIR_ERR_ARG;
  if (P <= 0 || S <= 0) return MBIR_ERR_SHAPE;
  int nb = (int)(((int64_t)P * S + kThreads - 1) / kThreads);
  if (nb > 8192) nb = 8192;
  vcd_positivity_kernel<<<nb, kThreads, 0, (hipStream_t)stream>>>(x, pix, P, S, alpha, delta);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int64_t mbir_vcd_apply_scratch(int32_t P, int32_t S) {
  if (P <= 0 || S <= 0) return 0;
  return (int64_t)apply_blocks(P) * (S + 1);
}

// Partial pairs times workgroups above which the finishing launch stops
// forming alpha in every workgroup (64 MiB of re-read partials, from L2).
constexpr int64_t kFoldAlphaPairs = 1 << 22;

// Update workgroups: >= 16 entries per lane, at most `cap` workgroups.
static int finish_update_blocks(int64_t n, int cap) {
  const int64_t nb = (n + 16LL * kThreads - 1) / (16LL * kThreads);
  return (int)(nb > cap ? cap : nb);
}

int32_t mbir_vcd_block_stats_rows(int32_t P) { return P > 0 ? apply_blocks(P) : 0; }

static int vcd_finish_any(void* e, const float* adelta, int64_t n_sino, float* x,
                          const int64_t* pix, int32_t P, int32_t S, const float* delta,
                          const double* dir_parts, int32_t n_dir, const double* fwd_parts,
                          int32_t n_fwd, const mbir_vcd_params* prm, float* alpha,
                          double* alpha_sum, double* block_stats, mbir_stream_t stream,
                          bool bf16) {
  if (!x || !pix || !delta || !alpha || !block_stats || n_sino < 0) return MBIR_ERR_ARG;
  if (n_sino > 0 && (!e || !adelta)) return MBIR_ERR_ARG;
  if (dir_parts && (!fwd_parts || !prm || !alpha_sum || n_dir < 0 || n_fwd < 0))
    return MBIR_ERR_ARG;
  if (P <= 0 || S <= 0) return MBIR_ERR_SHAPE;
  FinishArgs a;
  a.e = e;
  a.adelta = adelta;
  a.n_sino = n_sino;
  a.nb_update = finish_update_blocks(n_sino, 1024);
  a.x = x;
  a.pix = pix;
  a.P = P;
  a.S = S;
  const int nb_a = apply_blocks(P);
  a.ppb = (P + nb_a - 1) / nb_a;
  a.delta = delta;
  a.alpha = alpha;
  a.alpha_sum = alpha_sum;
  a.dir_parts = dir_parts;
  a.n_dir = n_dir;
  a.fwd_parts = fwd_parts;
  a.n_fwd = n_fwd;
  a.fm = prm ? (double)prm->fm_constant : 0.0;
  a.max_alpha = prm ? (double)prm->max_alpha : 0.0;
  a.part = block_stats;
  hipStream_t s = (hipStream_t)stream;
  const int64_t grid = (int64_t)a.nb_update + nb_a;
  if (dir_parts && (int64_t)(n_dir + n_fwd) * grid > kFoldAlphaPairs) {
    // Every workgroup would re-read too many partials: one workgroup forms
    // alpha first (the same arithmetic), the finishing launch reads it with
    // more update workgroups.
    vcd_alpha_kernel<<<1, kThreads, 0, s>>>(dir_parts, n_dir, fwd_parts, n_fwd, a.fm,
                                            a.max_alpha, alpha, alpha_sum);
    MBIR_CHECK_LAUNCH();
    a.dir_parts = nullptr;
    a.nb_update = finish_update_blocks(n_sino, 8192);
  }
  const size_t wd = bf16 ? 8 : 16;  // bytes of four entries
  const bool vec = n_sino % 4 == 0 && reinterpret_cast<uintptr_t>(e) % wd == 0 &&
                   reinterpret_cast<uintptr_t>(adelta) % 16 == 0;
  const int nb = a.nb_update + nb_a;
  if (bf16) {
    if (vec) vcd_finish_kernel<bf16_bits, true><<<nb, kThreads, 0, s>>>(a);
    else vcd_finish_kernel<bf16_bits, false><<<nb, kThreads, 0, s>>>(a);
  } else {
    if (vec) vcd_finish_kernel<float, true><<<nb, kThreads, 0, s>>>(a);
    else vcd_finish_kernel<float, false><<<nb, kThreads, 0, s>>>(a);
  }
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int mbir_vcd_finish(float* error_vcr, const float* adelta_vcr, int64_t n_sino, float* x,
                    const int64_t* pix, int32_t P, int32_t S, const float* delta,
                    const double* dir_parts, int32_t n_dir, const double* fwd_parts,
                    int32_t n_fwd, const mbir_vcd_params* prm, float* alpha, double* alpha_sum,
                    double* block_stats, mbir_stream_t stream) {
  return vcd_finish_any(error_vcr, adelta_vcr, n_sino, x, pix, P, S, delta, dir_parts, n_dir,
                        fwd_parts, n_fwd, prm, alpha, alpha_sum, block_stats, stream, false);
}

int mbir_vcd_finish_bf16(uint16_t* error_vcr, const float* adelta_vcr, int64_t n_sino, float* x,
                         const int64_t* pix, int32_t P, int32_t S, const float* delta,
                         const double* dir_parts, int32_t n_dir, const double* fwd_parts,
                         int32_t n_fwd, const mbir_vcd_params* prm, float* alpha,
                         double* alpha_sum, double* block_stats, mbir_stream_t stream) {
  return vcd_finish_any(error_vcr, adelta_vcr, n_sino, x, pix, P, S, delta, dir_parts, n_dir,
                        fwd_parts, n_fwd, prm, alpha, alpha_sum, block_stats, stream, true);
}

int mbir_vcd_block_stats_reduce(double* block_stats, int32_t rows, int32_t S, double* acc,
                                mbir_stream_t stream) {
  if (!block_stats || !acc || rows < 0) return MBIR_ERR_ARG;
  if (S <= 0) return MBIR_ERR_SHAPE;
  if (rows == 0) return MBIR_OK;
  vcd_stats_final_kernel<true><<<S + 1, kThreads, 0, (hipStream_t)stream>>>(block_stats, rows,
                                                                             S, acc);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int mbir_vcd_apply(float* x, const int64_t* pix, int32_t P, int32_t S, const float* delta,
                   const float* alpha, double* acc, double* scratch, mbir_stream_t stream) {
  if (!x || !pix || !delta || !alpha || !acc || !scratch) return MBIR_ERR_ARG;
  if (P <= 0 || S <= 0) return MBIR_ERR_SHAPE;
  hipStream_t s = (hipStream_t)stream;
  const int nb = apply_blocks(P);
  const int ppb = (P + nb - 1) / nb;
  vcd_apply_kernel<<<nb, kThreads, 0, s>>>(x, pix, P, S, delta, alpha, ppb, scratch);
  MBIR_CHECK_LAUNCH();
  vcd_stats_final_kernel<false><<<S + 1, kThreads, 0, s>>>(scratch, nb, S, acc);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int32_t mbir_reduce_blocks(int64_t n) { return red_blocks(n); }

int mbir_sumsq_f64(const float* a, const float* w, int64_t n, double* scratch, double* out,
                   mbir_stream_t stream) {
  if (!a || !scratch || !out || n < 0) return MBIR_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int nb = red_blocks(n);
  sumsq_partial_kernel<float><<<nb, kThreads, 0, s>>>(a, w, n, scratch);
  MBIR_CHECK_LAUNCH();
  final3_kernel<<<1, kThreads, 0, s>>>(scratch, nb, 3, out);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int mbir_sumsq_bf16(const uint16_t* a, const uint16_t* w, int64_t n, double* scratch,
                    double* out, mbir_stream_t stream) {
  if (!a || !scratch || !out || n < 0) return MBIR_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int nb = red_blocks(n);
  sumsq_partial_kernel<bf16_bits><<<nb, kThreads, 0, s>>>(a, w, n, scratch);
  MBIR_CHECK_LAUNCH();
  final3_kernel<<<1, kThreads, 0, s>>>(scratch, nb, 3, out);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

static int error_update_grid(int64_t n) {
  int64_t nb = (n + kThreads - 1) / kThreads;
  return (int)(nb > 16384 ? 16384 : (nb < 1 ? 1 : nb));
}

int mbir_vcd_error_update(float* error_vcr, const float* adelta_vcr, const float* alpha,
                          int64_t n, mbir_stream_t stream) {
  if (!error_vcr || !adelta_vcr || !alpha || n < 0) return MBIR_ERR_ARG;
  if (n == 0) return MBIR_OK;
  error_update_kernel<float><<<error_update_grid(n), kThreads, 0, (hipStream_t)stream>>>(
      error_vcr, adelta_vcr, alpha, n);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int mbir_vcd_error_update_bf16(uint16_t* error_vcr, const float* adelta_vcr, const float* alpha,
                               int64_t n, mbir_stream_t stream) {
  if (!error_vcr || !adelta_vcr || !alpha || n < 0) return MBIR_ERR_ARG;
  if (n == 0) return MBIR_OK;
  error_update_kernel<bf16_bits><<<error_update_grid(n), kThreads, 0, (hipStream_t)stream>>>(
      error_vcr, adelta_vcr, alpha, n);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int mbir_f32_to_bf16(const float* in, uint16_t* out, int64_t n, mbir_stream_t stream) {
  if (!in || !out || n < 0) return MBIR_ERR_ARG;
  if (n == 0) return MBIR_OK;
  int nb = (int)((n + kThreads - 1) / kThreads);
  if (nb > 8192) nb = 8192;
  convert_kernel<float, bf16_bits><<<nb, kThreads, 0, (hipStream_t)stream>>>(in, out, n);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int mbir_bf16_to_f32(const uint16_t* in, float* out, int64_t n, mbir_stream_t stream) {
  if (!in || !out || n < 0) return MBIR_ERR_ARG;
  if (n == 0) return MBIR_OK;
  int nb = (int)((n + kThreads - 1) / kThreads);
  if (nb > 8192) nb = 8192;
  convert_kernel<bf16_bits, float><<<nb, kThreads, 0, (hipStream_t)stream>>>(in, out, n);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int mbir_sumabs_f64(const float* a, int64_t n, double* scratch, double* out,
                    mbir_stream_t stream) {
  if (!a || !scratch || !out || n < 0) return MBIR_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int nb = red_blocks(n);
  abs_dot_partial_kernel<<<nb, kThreads, 0, s>>>(a, nullptr, n, scratch);
  MBIR_CHECK_LAUNCH();
  final3_kernel<<<1, kThreads, 0, s>>>(scratch, nb, 1, out);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

int mbir_dot_f64(const float* a, const float* b, int64_t n, double* scratch, double* out,
                 mbir_stream_t stream) {
  if (!a || !b || !scratch || !out || n < 0) return MBIR_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int nb = red_blocks(n);
  abs_dot_partial_kernel<<<nb, kThreads, 0, s>>>(a, b, n, scratch);
  MBIR_CHECK_LAUNCH();
  final3_kernel<<<1, kThreads, 0, s>>>(scratch, nb, 1, out);
  MBIR_CHECK_LAUNCH();
  return MBIR_OK;
}

}  // extern "C"
