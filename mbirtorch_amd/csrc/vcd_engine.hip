// Native VCD schedule runner: the per-subset launch sequence of the VCD
// engine, issued from C++ on one stream with no host synchronisation.
//
// One subset step (tomography_model.py:2508-2657, vcd_subset_updater):
// weighted back projection of the error sinogram at the subset (gradient),
// prior + direction (qGGMRF surrogate, proximal prior or cone DC damping),
// forward projection of the direction fused with the FP64 line-search sums,
// the on-device step size alpha, optional positivity clip, the error update
// e -= alpha*A(delta), and the recon update with its per-slice statistics.
// The subset order comes from the caller (the reference's global-RNG
// permutation, drawn on the host).
#include "common.h"

#include <algorithm>

using namespace mbir;

namespace {

int check_job(const mbir_vcd_job* j) {
  if (!j) return MBIR_ERR_ARG;
  if (j->geometry == MBIR_GEOMETRY_PARALLEL && !j->par_geom) return MBIR_ERR_ARG;
  if (j->geometry == MBIR_GEOMETRY_CONE && !j->cone_geom) return MBIR_ERR_ARG;
  if (j->geometry != MBIR_GEOMETRY_PARALLEL && j->geometry != MBIR_GEOMETRY_CONE)
    return MBIR_ERR_ARG;
  if (!j->view_table || !j->flat_recon || !j->error_vcr || !j->hessian || !j->subset_pixels ||
      !j->subset_sizes || !j->back_parts || !j->delta || !j->dir_parts || !j->fwd_parts ||
      !j->alpha || !j->apply_scratch || !j->red_scratch)
    return MBIR_ERR_ARG;
  if (j->dc_profile && !j->dc_scratch) return MBIR_ERR_ARG;
  if (j->sino_dtype != MBIR_DTYPE_F32 && j->sino_dtype != MBIR_DTYPE_BF16) return MBIR_ERR_ARG;
  if (j->prm.num_slices <= 0 || j->num_subsets <= 0) return MBIR_ERR_SHAPE;
  return MBIR_OK;
}

#define MBIR_TRY(x)          \
  do {                       \
    const int st_ = (x);     \
    if (st_ != MBIR_OK) return st_; \
  } while (0)

int subset_step(const mbir_vcd_job* j, int k, double* stats, hipStream_t s) {
  const int S = j->prm.num_slices;
  const int P = j->subset_sizes[k];
  const int64_t* pix = j->subset_pixels[k];
  const int32_t* runs = j->subset_runs ? j->subset_runs[k] : nullptr;
  const bool par = j->geometry == MBIR_GEOMETRY_PARALLEL;
  const bool h = j->sino_dtype == MBIR_DTYPE_BF16;
  // bfloat16 storage: the error sinogram and weights are 16-bit words.
  const uint16_t* e16 = reinterpret_cast<const uint16_t*>(j->error_vcr);
  uint16_t* e16w = reinterpret_cast<uint16_t*>(j->error_vcr);
  const uint16_t* w16 = reinterpret_cast<const uint16_t*>(j->weights_vcr);
  if (P <= 0) return MBIR_OK;
  if (h) runs = nullptr;  // the binned forward form serves bfloat16 storage
  const int ns = par ? mbir_par_back_splits(j->par_geom, P) : mbir_cone_back_splits(j->cone_geom, P);
  if ((int64_t)ns * P * S > j->back_parts_capacity) return MBIR_ERR_WORKSPACE;
  if (par && h)
    MBIR_TRY(mbir_par_back_bf16(e16, w16, pix, P, j->view_table, j->par_geom, ns, j->back_parts,
                                s));
  else if (par)
    MBIR_TRY(mbir_par_back(j->error_vcr, j->weights_vcr, pix, P, j->view_table, j->par_geom, 1,
                           ns, j->back_parts, s));
  else if (h)
    MBIR_TRY(mbir_cone_back_bf16(e16, w16, pix, P, j->view_table, j->cone_geom, ns,
                                 j->back_parts, s));
  else
    MBIR_TRY(mbir_cone_back(j->error_vcr, j->weights_vcr, pix, P, j->view_table, j->cone_geom,
                            1, ns, j->back_parts, s));
  const int n_dir = mbir_vcd_direction_partials_count(P, S);
  if (j->dc_profile)
    MBIR_TRY(mbir_vcd_direction_dc(j->flat_recon, pix, P, j->back_parts, ns, j->hessian, j->prox,
                                   &j->prm, j->dc_profile, j->dc_scratch, j->delta, j->dir_parts,
                                   s));
  else
    MBIR_TRY(mbir_vcd_direction(j->flat_recon, pix, P, j->back_parts, ns, j->hessian, j->prox,
                                &j->prm, j->delta, j->dir_parts, s));
  // The line search forms only sum w (A delta)^2: the linear term
  // <W e, A delta> = <A^T W e, delta> came with the direction partials from
  // the back projection, so the error sinogram is not read here.
  int n_fwd;
  float* adelta = j->positivity ? nullptr : j->adelta;
  if (par) {
    if (h)
      MBIR_TRY(mbir_par_fwd_linesearch_bf16(j->delta, pix, P, j->view_table, j->par_geom,
                                            nullptr, w16, adelta, j->fwd_parts, s));
    else
      MBIR_TRY(mbir_par_fwd_linesearch(j->delta, pix, P, runs, j->view_table, j->par_geom,
                                       nullptr, j->weights_vcr, adelta, j->fwd_parts, s));
    n_fwd = mbir_par_fwd_partials_count(j->par_geom, runs != nullptr);
  } else {
    if (h)
      MBIR_TRY(mbir_cone_fwd_linesearch_bf16(j->delta, pix, P, j->view_table, j->cone_geom,
                                             nullptr, w16, adelta, j->fwd_parts, s));
    else
      MBIR_TRY(mbir_cone_fwd_linesearch(j->delta, pix, P, j->view_table, j->cone_geom, nullptr,
                                        j->weights_vcr, adelta, j->fwd_parts, s));
    n_fwd = mbir_cone_fwd_partials_count(j->cone_geom);
  }
  // The positivity clip changes delta after the line search, so A(delta)
  // is re-projected then; otherwise the stored A(delta) is streamed by the
  // finishing launch, which also forms alpha.
  double* block_stats = j->apply_scratch;
  if (j->adelta && !j->positivity) {
    if (h)
      return mbir_vcd_finish_bf16(e16w, j->adelta, j->sino_size, j->flat_recon, pix, P, S,
                                  j->delta, j->dir_parts, n_dir, j->fwd_parts, n_fwd, &j->prm,
                                  j->alpha, stats + S + 1, block_stats, s);
    return mbir_vcd_finish(j->error_vcr, j->adelta, j->sino_size, j->flat_recon, pix, P, S,
                           j->delta, j->dir_parts, n_dir, j->fwd_parts, n_fwd, &j->prm, j->alpha,
                           stats + S + 1, block_stats, s);
  }
  MBIR_TRY(mbir_vcd_alpha(j->dir_parts, n_dir, j->fwd_parts, n_fwd, &j->prm, j->alpha,
                          stats + S + 1, s));
  if (j->positivity) MBIR_TRY(mbir_vcd_positivity(j->flat_recon, pix, P, S, j->alpha, j->delta, s));
  if (par && h)
    MBIR_TRY(mbir_par_fwd_update_bf16(j->delta, pix, P, j->view_table, j->par_geom, j->alpha,
                                      e16w, s));
  else if (par)
    MBIR_TRY(mbir_par_fwd_update(j->delta, pix, P, runs, j->view_table, j->par_geom, j->alpha,
                                 j->error_vcr, s));
  else if (h)
    MBIR_TRY(mbir_cone_fwd_update_bf16(j->delta, pix, P, j->view_table, j->cone_geom, j->alpha,
                                       e16w, s));
  else
    MBIR_TRY(mbir_cone_fwd_update(j->delta, pix, P, j->view_table, j->cone_geom, j->alpha,
                                  j->error_vcr, s));
  return mbir_vcd_finish(nullptr, nullptr, 0, j->flat_recon, pix, P, S, j->delta, nullptr, 0,
                         nullptr, 0, nullptr, j->alpha, nullptr, block_stats, s);
}

}  // namespace

extern "C" {

int mbir_vcd_run(const mbir_vcd_job* job, const int32_t* subsets, int32_t num_steps,
                 double* stats, int32_t end_iteration, mbir_stream_t stream) {
  MBIR_TRY(check_job(job));
  if (num_steps < 0 || (num_steps > 0 && !subsets) || !stats) return MBIR_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < num_steps; ++i) {
    const int k = subsets[i];
    if (k < 0 || k >= job->num_subsets) return MBIR_ERR_ARG;
    MBIR_TRY(subset_step(job, k, stats, s));
  }
  if (end_iteration) {
    const int S = job->prm.num_slices;
    int rows = 0;
    for (int k = 0; k < job->num_subsets; ++k)
      rows = std::max(rows, (int)mbir_vcd_block_stats_rows(job->subset_sizes[k]));
    MBIR_TRY(mbir_vcd_block_stats_reduce(job->apply_scratch, rows, S, stats, s));
    if (job->sino_dtype == MBIR_DTYPE_BF16)
      MBIR_TRY(mbir_sumsq_bf16(reinterpret_cast<const uint16_t*>(job->error_vcr),
                               reinterpret_cast<const uint16_t*>(job->weights_vcr),
                               job->sino_size, job->red_scratch, stats + S + 2, s));
    else
      MBIR_TRY(mbir_sumsq_f64(job->error_vcr, job->weights_vcr, job->sino_size,
                              job->red_scratch, stats + S + 2, s));
    MBIR_TRY(mbir_sumabs_f64(job->flat_recon, job->recon_size, job->red_scratch, stats + S + 5,
                             s));
  }
  return MBIR_OK;
}

}  // extern "C"
