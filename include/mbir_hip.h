/*
 * mbir_hip.h -- C-ABI of the MI355X (gfx950) projector / VCD library
 * (libmbir_hip.so).
 *
 * Conventions
 *  - Every pointer is a device pointer owned by the caller; the library never
 *    allocates.  Scratch arrays (partials) are passed in, sized by the
 *    *_count / *_bytes queries below.
 *  - Every call is asynchronous on `stream` (a hipStream_t) for the current
 *    device; the caller sets the device.
 *  - Sinogram-like arrays are float32.  `layout` MBIR_LAYOUT_VRC is the
 *    reference's (views, rows, channels); MBIR_LAYOUT_VCR is channel-major
 *    (views, channels, rows), the library's internal layout (the reference
 *    bodies transpose to it on every call: horizontal_fan.py:38-43,
 *    triton_parallel.py:314-330).
 *  - Volumes are flat (rows*cols, slices), the reference's flat_recon
 *    (tomography_model.py:1714-1760); voxel cylinders at a pixel list are
 *    (num_pixels, slices).
 *  - Return 0 on success, a negative MBIR_ERR_* code otherwise; no call
 *    touches memory after a non-zero return.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference checkout).
 */
#ifndef MBIR_HIP_H
#define MBIR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBIR_OK 0
#define MBIR_ERR_SHAPE (-1)
#define MBIR_ERR_GEOMETRY (-2)
#define MBIR_ERR_LAUNCH (-3)
#define MBIR_ERR_WORKSPACE (-4)
#define MBIR_ERR_ARG (-5)

#define MBIR_LAYOUT_VRC 0
#define MBIR_LAYOUT_VCR 1

typedef void* mbir_stream_t; /* hipStream_t */

/* Parallel-beam geometry: the scalars of _view_batch_args
 * (parallel_beam.py:278-291) plus derived constants.  Per-view constants
 * (cos, sin, W, ws, (W+1)/2, min(W,1), 0, 0) come in a separate float32
 * table of 8 floats per view, computed on the host with the reference's
 * float32 operation order (parallel_beam.py:38-51). */
typedef struct {
  int32_t num_rows;      /* recon rows */
  int32_t num_cols;      /* recon columns */
  int32_t num_slices;    /* recon slices == detector rows */
  int32_t num_channels;  /* detector channels */
  int32_t num_views;
  int32_t psf_radius;    /* reference tap radius (parallel_beam.py:172-180) */
  int32_t two_tap;       /* 1 when every view has W <= 1: the two taps
                            floor(n_p), floor(n_p)+1 hold every nonzero weight */
  int32_t unit_channel;  /* 1 when delta_det_channel == 1 (division is exact) */
  float delta_det_channel;
  float det_channel_offset;
  float delta_voxel;
  float delta_voxel_row;
  float det_center;      /* (C - 1) / 2 */
  float row_center;      /* (num_rows - 1) / 2 */
  float col_center;      /* (num_cols - 1) / 2 */
  float pad;
} mbir_par_geom;

/* Cone-beam geometry: the scalars of ConeBeamModel._view_batch_args
 * (cone_beam.py:397-414) plus derived float32 constants.  Per-view constants
 * (cos, sin, angle, z_offset = recon_slice_offset - z_shift) come in a float32
 * table of 4 floats per view. */
typedef struct {
  int32_t num_rows;      /* recon rows */
  int32_t num_cols;      /* recon columns */
  int32_t num_slices;    /* recon slices */
  int32_t num_det_rows;
  int32_t num_channels;
  int32_t num_views;
  int32_t psf_radius;    /* channel taps and back vertical taps (cone_beam.py:519-553) */
  int32_t bp_psf_radius; /* forward vertical taps */
  int32_t curved;        /* 1: cylindrical detector of radius SDD */
  int32_t unit_channel;  /* delta_det_channel == 1 */
  int32_t pad0, pad1;
  float delta_voxel, delta_voxel_row, delta_voxel_slice;
  float delta_det_channel, delta_det_row;
  float det_channel_offset, det_row_offset;
  float source_detector_dist;  /* may be +inf */
  float source_iso_dist;       /* float32(SDD / M) */
  float inv_magnification;     /* float32(1 / M) */
  float det_center_channel;    /* (C - 1) / 2 */
  float det_center_row;        /* (R - 1) / 2 */
  float row_center, col_center, slice_center; /* (N - 1) / 2 of the recon axes */
  float half_slice_span;       /* float32(delta_voxel_slice * (S - 1) / 2) */
  float area;                  /* float32(delta_voxel_row * delta_voxel) */
  float pad2;
} mbir_cone_geom;

/* qGGMRF / VCD scalars (qggmrf.py:22-74, tomography_model.py:2463-2472). */
typedef struct {
  float b[6];            /* row+, row-, col+, col-, slice+, slice- (normalised) */
  float two_b[6];        /* float32(b[i] * 2.0) as the reference forms it */
  float sigma_x, p, q, T;
  float a_min;           /* float32(T * sigma_x * eps32) */
  float t_sigma;         /* float32(T * sigma_x) */
  float q_minus_p, p_minus_2, q_over_p;
  float two_sigma_p;     /* float32(2 * sigma_x ** p) */
  float fm_constant;     /* float32(1 / sigma_y**2) */
  float neg_fm_constant; /* float32(-(1 / sigma_y**2)) */
  float inv_sigma_prox2; /* float32(1 / sigma_prox**2) (proximal-map prior) */
  float max_alpha;
  int32_t num_rows, num_cols, num_slices;
  int32_t use_prox;      /* 1: proximal-map prior instead of qGGMRF */
} mbir_vcd_params;

/* ---- library info ------------------------------------------------------ */
int mbir_version(void);

/* ---- parallel-beam projectors ----------------------------------------- */

/* Sparse forward projection of voxel cylinders values (num_pixels, slices)
 * at pixel_indices into a full sinogram, written (not accumulated) in
 * `layout`.  `row_runs` (optional, 3 ints per image row: list position of
 * the row's first listed pixel, its column, the run length; length 0 for
 * rows without listed pixels) declares that every image row's pixels form
 * one run of consecutive columns in list order; the ray-driven gather is
 * then used, otherwise the binned gather.  When num_slices % 4 == 0 the
 * row arrays (values, sinogram, weights) must be 16-byte aligned
 * (MBIR_ERR_ARG otherwise).  Replaces the forward body
 * _parallel_forward_view_batch (parallel_beam.py:56-73; Triton body
 * triton_parallel.py:598-745) driven over the whole view range by
 * Projectors.sparse_forward_project_view_range (projectors.py:518-588). */
int mbir_par_forward(const float* values, const int64_t* pixel_indices,
                     int32_t num_pixels, const int32_t* row_runs,
                     const float* view_table, const mbir_par_geom* geom,
                     int32_t layout, float* sino_out, mbir_stream_t stream);

/* Number of view splits mbir_par_back uses for num_pixels (>= 1). */
int32_t mbir_par_back_splits(const mbir_par_geom* geom, int32_t num_pixels);

/* Sparse back projection of a channel-major sinogram (optionally multiplied
 * by weights of the same layout) onto voxel cylinders, with the weights
 * raised to coeff_power (1, or 2 for the Hessian diagonal).  Writes
 * num_splits partial sums (num_splits, num_pixels, slices) over disjoint view
 * ranges into `partials`; num_splits == 1 writes the result itself.
 * Replaces _parallel_back_view_batch (parallel_beam.py:76-94; Triton body
 * triton_parallel.py:250-369) driven by
 * Projectors.sparse_back_project_view_range (projectors.py:590-646). */
int mbir_par_back(const float* sino_vcr, const float* weights_vcr,
                  const int64_t* pixel_indices, int32_t num_pixels,
                  const float* view_table, const mbir_par_geom* geom,
                  int32_t coeff_power, int32_t num_splits, float* partials,
                  mbir_stream_t stream);

/* out[i] = sum_k parts[k * n + i], k ascending (deterministic). */
int mbir_reduce_splits(const float* parts, int32_t num_splits, int64_t n,
                       float* out, mbir_stream_t stream);

/* Number of FP64 partial pairs mbir_par_fwd_linesearch writes: with
 * row_runs given (ray != 0) or without (ray == 0). */
int32_t mbir_par_fwd_partials_count(const mbir_par_geom* geom, int32_t ray);

/* Forward projection of the subset direction delta fused with the line
 * search reductions lin = sum(w * e * A delta), quad = sum(w * (A delta)^2)
 * (w == NULL means unit weights), FP64, one (lin, quad) pair per workgroup
 * into partials[2 * mbir_par_fwd_partials_count(geom, row_runs != NULL)].
 * error_vcr == NULL leaves lin at 0 and reads no error sinogram: the caller
 * takes <W e, A delta> = <A^T W e, delta> from the back projection (the
 * third value of each mbir_vcd_direction partial).  `row_runs` as for
 * mbir_par_forward.  With adelta_vcr non-NULL, A delta is also stored there
 * (channel-major, float32) for mbir_vcd_error_update.  Replaces
 * sparse_forward_project(delta) + _forward_lin_quad_const/_weighted
 * (tomography_model.py:2592-2610, 65-74). */
int mbir_par_fwd_linesearch(const float* delta, const int64_t* pixel_indices,
                            int32_t num_pixels, const int32_t* row_runs,
                            const float* view_table, const mbir_par_geom* geom,
                            const float* error_vcr, const float* weights_vcr,
                            float* adelta_vcr, double* partials, mbir_stream_t stream);

/* error -= alpha[0] * (A delta), recomputing A delta in registers (no
 * sinogram-sized transient).  `row_runs` as for mbir_par_forward.  Replaces
 * error_sinogram.sub_(alpha * delta_sinogram) in _apply_update
 * (tomography_model.py:77-87). */
int mbir_par_fwd_update(const float* delta, const int64_t* pixel_indices,
                        int32_t num_pixels, const int32_t* row_runs,
                        const float* view_table, const mbir_par_geom* geom,
                        const float* alpha, float* error_vcr, mbir_stream_t stream);

/* bfloat16 storage (raw bits, uint16_t) of the channel-major error sinogram
 * and weights, with float32 arithmetic: the VCD back projection (coeff_power
 * 1) and the binned forward line search / update.  Same contracts as the
 * float32 entry points (line-search partials: mbir_par_fwd_partials_count(
 * geom, 0)); error_vcr is rounded to nearest even on update.  No reference
 * counterpart: the storage format of the mixed-precision path. */
int mbir_par_back_bf16(const uint16_t* sino_vcr, const uint16_t* weights_vcr,
                       const int64_t* pixel_indices, int32_t num_pixels,
                       const float* view_table, const mbir_par_geom* geom,
                       int32_t num_splits, float* partials, mbir_stream_t stream);
int mbir_par_fwd_linesearch_bf16(const float* delta, const int64_t* pixel_indices,
                                 int32_t num_pixels, const float* view_table,
                                 const mbir_par_geom* geom, const uint16_t* error_vcr,
                                 const uint16_t* weights_vcr, float* adelta_vcr,
                                 double* partials, mbir_stream_t stream);
int mbir_par_fwd_update_bf16(const float* delta, const int64_t* pixel_indices,
                             int32_t num_pixels, const float* view_table,
                             const mbir_par_geom* geom, const float* alpha,
                             uint16_t* error_vcr, mbir_stream_t stream);

/* ---- cone-beam projectors ---------------------------------------------- */

/* Sparse cone forward projection (detector-side vertical fan, then the
 * horizontal fan), written in `layout`.  Replaces _cone_forward_view_batch
 * (cone_beam.py:161-234; Triton body triton_cone.py:640-746). */
int mbir_cone_forward(const float* values, const int64_t* pixel_indices,
                      int32_t num_pixels, const float* view_table,
                      const mbir_cone_geom* geom, int32_t layout, float* sino_out,
                      mbir_stream_t stream);

/* Number of view splits mbir_cone_back uses for num_pixels (>= 1). */
int32_t mbir_cone_back_splits(const mbir_cone_geom* geom, int32_t num_pixels);

/* Sparse cone back projection of a channel-major sinogram (optionally times
 * weights), weights raised to coeff_power, as num_splits partial sums over
 * view ranges (num_splits, num_pixels, slices).  Replaces
 * _cone_back_view_batch (cone_beam.py:237-299; Triton body
 * triton_cone.py:333-462). */
int mbir_cone_back(const float* sino_vcr, const float* weights_vcr,
                   const int64_t* pixel_indices, int32_t num_pixels,
                   const float* view_table, const mbir_cone_geom* geom,
                   int32_t coeff_power, int32_t num_splits, float* partials,
                   mbir_stream_t stream);

/* Number of FP64 partial pairs the cone forward line search writes. */
int32_t mbir_cone_fwd_partials_count(const mbir_cone_geom* geom);

/* Cone forward of delta fused with the line-search reductions (see
 * mbir_par_fwd_linesearch). */
int mbir_cone_fwd_linesearch(const float* delta, const int64_t* pixel_indices,
                             int32_t num_pixels, const float* view_table,
                             const mbir_cone_geom* geom, const float* error_vcr,
                             const float* weights_vcr, float* adelta_vcr, double* partials,
                             mbir_stream_t stream);

/* error -= alpha[0] * (A delta) for the cone operator. */
int mbir_cone_fwd_update(const float* delta, const int64_t* pixel_indices,
                         int32_t num_pixels, const float* view_table,
                         const mbir_cone_geom* geom, const float* alpha,
                         float* error_vcr, mbir_stream_t stream);

/* ---- VCD subset step --------------------------------------------------- */

/* Number of FP64 partial triples mbir_vcd_direction (and _dc) writes.  Not
 * monotonic in num_pixels: small subsets spread each element over more lanes
 * and so over more workgroups. */
int32_t mbir_vcd_direction_partials_count(int32_t num_pixels,
                                          int32_t num_slices);

/* Prior gradient/Hessian at the subset (qGGMRF surrogate, or the proximal
 * prior when prm->use_prox), forward gradient -fm*sum(back partials) and
 * Hessian fm*fm_hessian[pix], the direction delta = -(g_f+g_p)/(H_f+H_p),
 * and FP64 partial triples per workgroup (sum g_p*delta, sum H_p*delta^2,
 * sum b*delta with b = sum of the back partials = A^T W e at the subset --
 * the line search's linear term, so the forward line search can skip the
 * error sinogram).  Replaces qggmrf_gradient_and_hessian_at_indices
 * (qggmrf.py:77-183), prox_gradient_at_indices (qggmrf.py:186-206),
 * _diagonal_update_direction and _prior_line_terms
 * (tomography_model.py:54-62, 2520-2576). */
int mbir_vcd_direction(const float* flat_recon, const int64_t* pixel_indices,
                       int32_t num_pixels, const float* back_parts,
                       int32_t num_splits, const float* fm_hessian,
                       const float* prox_input, const mbir_vcd_params* prm,
                       float* delta, double* partials, mbir_stream_t stream);

/* Scratch floats mbir_vcd_direction_dc needs: 5 * num_pixels * num_slices
 * plus the per-slice means. */
int64_t mbir_vcd_direction_dc_scratch(int32_t num_pixels, int32_t num_slices);

/* Cone DC-damped direction (cone_beam.py:48-55, 476-496):
 * g = g_f + g_p, h_inv = 1/(H_f + H_p), gbar_k = sum_p h_inv*g / sum_p h_inv
 * per slice over the subset (FP64, one workgroup per slice),
 * delta = -(g - (1 - s_k) * gbar_k) * h_inv, with the prior line-term partials
 * as mbir_vcd_direction writes them.  slice_profile holds s_k (float32). */
int mbir_vcd_direction_dc(const float* flat_recon, const int64_t* pixel_indices,
                          int32_t num_pixels, const float* back_parts,
                          int32_t num_splits, const float* fm_hessian,
                          const float* prox_input, const mbir_vcd_params* prm,
                          const float* slice_profile, float* scratch, float* delta,
                          double* partials, mbir_stream_t stream);

/* alpha = clamp((fm*lin - prior_lin) / (fm*quad + prior_quad + eps32),
 * eps32, max_alpha), reduced in a fixed order from the direction triples
 * (prior_lin, prior_quad, back_lin) and the line-search pairs (fwd_lin,
 * quad), with lin = fwd_lin + back_lin (exactly one of the two is nonzero: a
 * line search run with the error sinogram must be paired with direction
 * triples whose third values are zero); writes alpha[0] (float32) and adds
 * it to alpha_sum[0] (FP64).  Replaces tomography_model.py:2612-2617. */
int mbir_vcd_alpha(const double* prior_parts, int32_t num_prior,
                   const double* fwd_parts, int32_t num_fwd,
                   const mbir_vcd_params* prm, float* alpha,
                   double* alpha_sum, mbir_stream_t stream);

/* delta = max(-x[pix] / (alpha + eps32), delta).  tomography_model.py:2624-2633. */
int mbir_vcd_positivity(const float* flat_recon, const int64_t* pixel_indices,
                        int32_t num_pixels, int32_t num_slices,
                        const float* alpha, float* delta, mbir_stream_t stream);

/* error -= alpha[0] * adelta over n elements (the A delta a line search
 * stored): the error update of _apply_update (tomography_model.py:77-87) as
 * one streaming pass.  The bf16 variant rounds the result to nearest even. */
int mbir_vcd_error_update(float* error_vcr, const float* adelta_vcr, const float* alpha,
                          int64_t n, mbir_stream_t stream);
int mbir_vcd_error_update_bf16(uint16_t* error_vcr, const float* adelta_vcr,
                               const float* alpha, int64_t n, mbir_stream_t stream);

/* Number of doubles of scratch mbir_vcd_apply needs. */
int64_t mbir_vcd_apply_scratch(int32_t num_pixels, int32_t num_slices);

/* x[pix] += alpha*delta, then acc[s] += sum_p (alpha*delta[p, s])^2 for
 * each slice s and acc[num_slices] += sum |alpha*delta|, FP64: per-workgroup
 * partials into `scratch`, then a fixed-order reduction (deterministic).
 * Replaces _apply_update's recon half and its per-slice / ell1 reductions
 * (tomography_model.py:77-87). */
int mbir_vcd_apply(float* flat_recon, const int64_t* pixel_indices,
                   int32_t num_pixels, int32_t num_slices, const float* delta,
                   const float* alpha, double* acc, double* scratch,
                   mbir_stream_t stream);

/* The end of a subset step as one launch: the step size, the error update
 * error -= alpha*adelta over n_sino elements (n_sino = 0: none), and
 * x[pix] += alpha*delta with its per-slice statistics.  With dir_parts set,
 * alpha is formed from the prior and line-search partials exactly as
 * mbir_vcd_alpha does (written to alpha[0], added to alpha_sum[0]);
 * dir_parts = NULL reads alpha[0].  The statistics -- sum_p (alpha*delta[p,
 * s])^2 per slice, then sum |alpha*delta| -- are ADDED into per-workgroup rows
 * block_stats[b*(num_slices+1) + c], b < mbir_vcd_block_stats_rows(P);
 * mbir_vcd_block_stats_reduce folds them into an accumulator and clears
 * them.  block_stats holds mbir_vcd_apply_scratch(max P, num_slices) doubles,
 * zero before its first use.  Fuses mbir_vcd_alpha, mbir_vcd_error_update and
 * mbir_vcd_apply (tomography_model.py:2612-2617, 77-87). */
int mbir_vcd_finish(float* error_vcr, const float* adelta_vcr, int64_t n_sino,
                    float* flat_recon, const int64_t* pixel_indices, int32_t num_pixels,
                    int32_t num_slices, const float* delta, const double* dir_parts,
                    int32_t num_dir, const double* fwd_parts, int32_t num_fwd,
                    const mbir_vcd_params* prm, float* alpha, double* alpha_sum,
                    double* block_stats, mbir_stream_t stream);
int mbir_vcd_finish_bf16(uint16_t* error_vcr, const float* adelta_vcr, int64_t n_sino,
                         float* flat_recon, const int64_t* pixel_indices,
                         int32_t num_pixels, int32_t num_slices, const float* delta,
                         const double* dir_parts, int32_t num_dir, const double* fwd_parts,
                         int32_t num_fwd, const mbir_vcd_params* prm, float* alpha,
                         double* alpha_sum, double* block_stats, mbir_stream_t stream);

/* Rows of block_stats that mbir_vcd_finish writes for a subset of P pixels. */
int32_t mbir_vcd_block_stats_rows(int32_t num_pixels);

/* acc[c] += sum_b block_stats[b*(num_slices+1) + c] for c <= num_slices in a
 * fixed order over rows b < rows, and clears those rows. */
int mbir_vcd_block_stats_reduce(double* block_stats, int32_t rows, int32_t num_slices,
                                double* acc, mbir_stream_t stream);

/* bfloat16-storage cone variants (see the parallel ones). */
int mbir_cone_back_bf16(const uint16_t* sino_vcr, const uint16_t* weights_vcr,
                        const int64_t* pixel_indices, int32_t num_pixels,
                        const float* view_table, const mbir_cone_geom* geom,
                        int32_t num_splits, float* partials, mbir_stream_t stream);
int mbir_cone_fwd_linesearch_bf16(const float* delta, const int64_t* pixel_indices,
                                  int32_t num_pixels, const float* view_table,
                                  const mbir_cone_geom* geom, const uint16_t* error_vcr,
                                  const uint16_t* weights_vcr, float* adelta_vcr,
                                  double* partials, mbir_stream_t stream);
int mbir_cone_fwd_update_bf16(const float* delta, const int64_t* pixel_indices,
                              int32_t num_pixels, const float* view_table,
                              const mbir_cone_geom* geom, const float* alpha,
                              uint16_t* error_vcr, mbir_stream_t stream);

/* ---- native VCD schedule runner ---------------------------------------- */

#define MBIR_GEOMETRY_PARALLEL 0
#define MBIR_GEOMETRY_CONE 1

#define MBIR_DTYPE_F32 0
#define MBIR_DTYPE_BF16 1

/* Everything one VCD reconstruction needs on the device, plus host tables of
 * its partition subsets (subset id -> device index array, size, optional
 * image-row run table).  Scratch sizes: back_parts >= back_parts_capacity
 * floats (max over subsets of splits * size * slices), delta >= max size *
 * slices floats, dir_parts >= 3 * the largest
 * mbir_vcd_direction_partials_count(size, slices) over the subset sizes
 * (small subsets use more workgroups: not always the largest subset's),
 * fwd_parts >= 2 * max forward partials count, apply_scratch >=
 * mbir_vcd_apply_scratch(max size, slices) doubles and zero before the first
 * run (it carries per-workgroup statistics to the end of each iteration,
 * which leaves it zero), red_scratch >= 3 * mbir_reduce_blocks(max(
 * sino_size, recon_size)), dc_scratch >= mbir_vcd_direction_dc_scratch when
 * dc_profile is set. */
typedef struct {
  int32_t geometry;                 /* MBIR_GEOMETRY_* */
  int32_t positivity;               /* clip x + alpha*delta at zero */
  int32_t num_subsets;              /* entries of the subset tables */
  int32_t sino_dtype;               /* MBIR_DTYPE_*: storage of error_vcr, weights_vcr */
  const mbir_par_geom* par_geom;    /* host pointer (parallel) */
  const mbir_cone_geom* cone_geom;  /* host pointer (cone) */
  const float* view_table;
  mbir_vcd_params prm;
  float* flat_recon;                /* (rows*cols, slices) */
  float* error_vcr;                 /* (views, channels, rows) */
  const float* weights_vcr;         /* NULL: unit weights */
  const float* hessian;             /* (rows*cols, slices) */
  const float* prox;                /* proximal-map input or NULL */
  const float* dc_profile;          /* cone DC-damping slice profile or NULL */
  const int64_t* const* subset_pixels;
  const int32_t* subset_sizes;
  const int32_t* const* subset_runs; /* NULL, or per subset NULL / run table */
  float* back_parts;
  int64_t back_parts_capacity;
  float* delta;
  double* dir_parts;
  double* fwd_parts;
  float* alpha;
  double* apply_scratch;
  float* dc_scratch;
  double* red_scratch;
  int64_t sino_size;                /* elements of error_vcr */
  int64_t recon_size;               /* elements of flat_recon */
  float* adelta;                    /* sino_size floats: the line search stores A delta
                                       and the error update streams it; NULL (or with
                                       positivity, which clips delta after the line
                                       search): the update re-projects delta */
} mbir_vcd_job;

/* Runs the subset steps subsets[0 .. num_steps) of one VCD iteration on
 * `stream`, accumulating into stats (num_slices + 6 doubles, zeroed by the
 * caller: per-slice sum (alpha*delta)^2, sum |alpha*delta|, sum alpha, then
 * the forward-model sums sum w*e^2, sum e^2, sum w and sum |x|).  The
 * per-slice sums and sum |alpha*delta| of all steps since the previous
 * iteration end, and the forward-model sums, land with end_iteration != 0;
 * sum alpha lands at every step.  No host synchronisation.  Replaces the subset loop of
 * vcd_partition_iterator / _vcd_recon (tomography_model.py:2662-2697,
 * 2851-2960). */
int mbir_vcd_run(const mbir_vcd_job* job, const int32_t* subsets, int32_t num_steps,
                 double* stats, int32_t end_iteration, mbir_stream_t stream);

/* ---- continuous-batching request engine -------------------------------- */

/* A native worker thread with `max_active` slots, each with its own HIP
 * stream (default priority) on the calling thread's current device.  Every
 * round it issues one subset step of every active job (mbir_vcd_run), so the
 * steps of different reconstructions overlap on the device; submitted jobs
 * wait in a queue until a slot frees.  No counterpart in the reference (it
 * has no request loop); the per-job work is exactly mbir_vcd_run's. */
typedef struct mbir_server mbir_server;

mbir_server* mbir_server_create(int32_t max_active);

/* Queues a job: schedule[iter_offsets[i] .. iter_offsets[i+1]) are the subset
 * ids of iteration i; stats holds num_iters blocks of num_slices + 6 doubles
 * (zeroed by the caller), one per iteration, filled as for mbir_vcd_run.
 * The job's device buffers and every host table it points to (subset
 * tables, schedule, iter_offsets) must stay valid until the job is returned
 * by mbir_server_poll.  The device work the job depends on must be complete
 * (or ordered before it) when it is submitted. */
int mbir_server_submit(mbir_server* server, const mbir_vcd_job* job,
                       const int32_t* schedule, const int32_t* iter_offsets,
                       int32_t num_iters, double* stats, int64_t tag);

/* Up to max_jobs finished jobs (tag, status), waiting up to timeout_ms for
 * the first; returns how many were written. */
int32_t mbir_server_poll(mbir_server* server, int64_t* tags, int32_t* status,
                         int32_t max_jobs, int32_t timeout_ms);

/* Jobs queued or running. */
int32_t mbir_server_in_flight(mbir_server* server);

/* Waits for every job, then stops the worker and releases the streams. */
void mbir_server_destroy(mbir_server* server);

/* ---- reductions -------------------------------------------------------- */

/* Scratch doubles the reductions need: 3 * mbir_reduce_blocks(n). */
int32_t mbir_reduce_blocks(int64_t n);

/* out[0] = sum w*a*a, out[1] = sum a*a, out[2] = sum w (w NULL: 1), FP64.
 * Backs get_forward_model_loss / _vcd_iteration_stats
 * (tomography_model.py:2323-2357). */
int mbir_sumsq_f64(const float* a, const float* w, int64_t n, double* scratch,
                   double* out, mbir_stream_t stream);

/* mbir_sumsq_f64 over bfloat16 arrays. */
int mbir_sumsq_bf16(const uint16_t* a, const uint16_t* w, int64_t n, double* scratch,
                    double* out, mbir_stream_t stream);

/* Storage conversion float32 <-> bfloat16 (round to nearest even). */
int mbir_f32_to_bf16(const float* in, uint16_t* out, int64_t n, mbir_stream_t stream);
int mbir_bf16_to_f32(const uint16_t* in, float* out, int64_t n, mbir_stream_t stream);

/* out[0] = sum |a|, FP64 (image_ell1, _memory_ledger.py:134-163). */
int mbir_sumabs_f64(const float* a, int64_t n, double* scratch, double* out,
                    mbir_stream_t stream);

/* out[0] = sum a*b, FP64 (adjoint inner products). */
int mbir_dot_f64(const float* a, const float* b, int64_t n, double* scratch,
                 double* out, mbir_stream_t stream);

/* ---- direct-reconstruction filter and layout -------------------------- */

/* Per-row 'valid' convolution with the 2C-1 taps (already scaled), with an
 * optional (rows, channels) pre-weight; input in VRC layout, output in
 * `out_layout`.  Replaces apply_row_filter (tomography_utils.py:29-78). */
int mbir_row_filter(const float* sino_vrc, const float* taps,
                    const float* row_weight, int32_t num_views,
                    int32_t num_rows, int32_t num_channels, int32_t out_layout,
                    float* out, mbir_stream_t stream);

/* VRC <-> VCR transpose; to_vcr != 0 converts VRC -> VCR. */
int mbir_transpose_sino(const float* in, int32_t num_views, int32_t num_rows,
                        int32_t num_channels, int32_t to_vcr, float* out,
                        mbir_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* MBIR_HIP_H */
