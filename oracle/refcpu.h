/*
 * refcpu — CPU restatement of texture-synthesis's per-pixel patch search.
 *
 * TEST INFRASTRUCTURE. This library is the correctness oracle and the timed
 * CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product library never links or calls it.
 *
 * Every function restates reference code; citations are to
 * EmbarkStudios/texture-synthesis (paths relative to the repo root).
 */
#ifndef REFCPU_H
#define REFCPU_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- PRNG layer (rand_pcg 0.3.1 / rand_core 0.6.3 / rand 0.8.5) ---- */
/* Lcg64Xsh32::new(state, stream) followed by n next_u32 draws. */
void orc_pcg_new_u32(uint64_t state, uint64_t stream, uint32_t n, uint32_t* out);
/* Pcg32::seed_from_u64(seed) followed by n next_u64 draws. */
void orc_pcg_seed_u64(uint64_t seed, uint32_t n, uint64_t* out);
/* Pcg32::from_seed(bytes) followed by n next_u64 draws. */
void orc_pcg_from_seed_u64(const uint8_t seed[16], uint32_t n, uint64_t* out);
/* Lcg64Xsh32::new(state, stream) then n gen_range::<u32>(lo..hi) draws. */
void orc_range_u32(uint64_t state, uint64_t stream, uint32_t lo, uint32_t hi, uint32_t n, uint32_t* out);
/* Lcg64Xsh32::new(state, stream) then n gen_range::<u64>(lo..hi) draws. */
void orc_range_u64(uint64_t state, uint64_t stream, uint64_t lo, uint64_t hi, uint32_t n, uint64_t* out);
/* Pcg32::seed_from_u64(seed).gen_range(0..len) as usize (ms.rs:386). */
uint64_t orc_pick_index(uint64_t seed, uint64_t len);

/* ---- cost tables (ms.rs:1109-1120, 1290-1311, 853-858) ---- */
/* full 256x256 tables as the reference builds them, for the given stage alpha. */
void orc_luts_full(float cauchy_dispersion, float alpha_s, int guided, float* my_cost65536, float* guide_cost65536);

/* ---- image resampling (image 0.23.12 imageops::resize, restated) ---- */
/* filter: 0 = Triangle, 1 = CatmullRom, 2 = Gaussian. RGBA8 in/out. */
int orc_resize(const uint8_t* src, uint32_t w, uint32_t h, uint32_t nw, uint32_t nh, int filter, uint8_t* dst);
/* Style-transfer guide steps (utils.rs:101-183): transform_to_guide_map
 * without a resize (blur sigma 2, grayscale) and match_histograms, in place. */
void orc_guide_map(const uint8_t* src, uint32_t w, uint32_t h, uint8_t* dst);
void orc_match_histograms(uint8_t* src, uint32_t w, uint32_t h, const uint8_t* tgt, uint32_t tw, uint32_t th);
/* ImagePyramid::new(img, Some(levels)): writes `levels` (>=1) full-size RGBA8 levels. */
int orc_pyramid(const uint8_t* src, uint32_t w, uint32_t h, uint32_t levels, uint8_t* out);

/* ---- synthesis job ---- */
typedef struct orc_params {
    uint32_t nearest_neighbors;        /* k */
    uint64_t random_sample_locations;  /* m */
    float cauchy_dispersion;
    float backtrack_percent;           /* p */
    uint32_t backtrack_stages;         /* P */
    float guide_alpha;
    uint64_t seed;
    uint32_t tiling_mode;
    uint64_t random_init;              /* UINT64_MAX = none */
} orc_params;

/* Jacobi batch schedule. px_per_batch > 0 forces fixed-size batches
 * (1 == the reference at max_thread_count(1)). Otherwise a batch inside the
 * redo segment holds max(1, (resolved - locked) >> redo_shift) items and a
 * batch inside the new-pixel segment max(1, (resolved - locked) >> new_shift)
 * (>> first_new_shift in the first stage when that is nonzero), capped at
 * max_batch, never crossing the segment boundary. */
typedef struct orc_sched {
    uint32_t px_per_batch;
    uint32_t redo_shift;
    uint32_t new_shift;
    uint32_t max_batch;
    uint32_t first_new_shift;
} orc_sched;

typedef struct orc_image { uint32_t w, h; const uint8_t* rgba; } orc_image;

typedef struct orc_job orc_job;

/* examples: finest-level RGBA8, already at their final size.
 * sample_kind[i]: 0 = All, 1 = Ignore, 2 = Image(sample_masks[i]).
 * example_guides / target_guide: NULL when unguided (all-or-none).
 * inpaint_mask: NULL unless inpainting; then inpaint_color is the example
 * at output size and inpaint_index its index in `examples`. */
orc_job* orc_job_new(const orc_params* prm, uint32_t n_examples, const orc_image* examples,
                     const int* sample_kind, const orc_image* sample_masks,
                     const orc_image* example_guides, const orc_image* target_guide,
                     const orc_image* inpaint_mask, const orc_image* inpaint_color,
                     uint32_t inpaint_index, uint32_t out_w, uint32_t out_h);
void orc_job_free(orc_job* j);
const char* orc_job_error(orc_job* j);

/* Runs Session::run (random init, then resolve) to completion. */
int orc_job_run(orc_job* j, const orc_sched* s, uint32_t n_threads);
/* Runs Session::run with the reference's own threading (ms.rs:815-1036),
 * for timing: per stage one worker while fewer than 1000 items are to be
 * redone, else n_threads workers on an atomic item counter, each item
 * reading and writing the shared maps at once, k nearest from the TreeGrid
 * (ms.rs:1313-1530; per-cell exact search, fan-out cells at the first
 * parallel stage). With more than one worker the result depends on thread
 * timing, as the reference's does; with one it equals orc_job_run with one
 * pixel per batch. */
int orc_job_run_reference_threads(orc_job* j, uint32_t n_threads);
/* Runs until the given stage index (0 = first stage, p_stage = P) and
 * work item, leaving the state frozen at that point. */
int orc_job_run_until(orc_job* j, const orc_sched* s, uint32_t n_threads, uint32_t stage_t, uint64_t item);

/* State download. Any pointer may be NULL. coord: (x,y,map) per px; ids: (patch,map). */
void orc_job_get(orc_job* j, uint8_t* rgba, uint32_t* coord, uint32_t* ids);
uint64_t orc_job_resolved(orc_job* j, uint32_t* flat, float* score, uint64_t cap);
uint64_t orc_job_locked(orc_job* j);
/* Generator::get_id_maps / get_uncertainty_map (ms.rs:603-653) of the
 * current state: RGBA8 w*h*4 each, any may be NULL. */
void orc_job_debug_maps(orc_job* j, uint8_t* patch_ids, uint8_t* map_ids, uint8_t* uncertainty);
/* per-stage: sum and count of the best scores of items resolved by pattern
 * matching with a numeric (non-NaN) cost, items done. */
uint32_t orc_job_stage_stats(orc_job* j, double* sums, uint64_t* counts, uint64_t* items);
/* resolved mask (1 byte per px) */
void orc_job_mask(orc_job* j, uint8_t* mask);

/* Frozen-state evaluation at the point orc_job_run_until stopped.
 * Takes the next n work items of the current stage (picking new pixels as
 * the reference would) and reports, per item, without committing anything:
 *   flat[i], nbr_count[i], nbrs[i*k + j] = (x,y) int32 pairs,
 *   weights[i*k + j] (f32), cand_count[i],
 *   cands[(i*(k+m) + c)*6] = (x, y, map, kind(0 coh,1 rand), patch, pmap),
 *   full_cost[i*(k+m)+c] = cost summed without early exit,
 *   best[i] = winning candidate index, best_score[i] = reference score.
 * The stage's LUTs and level are those of the current stage. Returns the
 * number of items described. */
uint64_t orc_job_describe(orc_job* j, uint64_t n, uint32_t* flat, uint32_t* nbr_count,
                          int32_t* nbrs, float* weights, uint32_t* cand_count, int32_t* cands,
                          float* full_cost, uint32_t* best, float* best_score);
/* current stage info: level, guided, alpha, stage seed, redo count, stage size, item cursor */
void orc_job_stage_info(orc_job* j, uint32_t* level, float* alpha_s, uint64_t* stage_seed,
                        uint64_t* redo, uint64_t* stage_size, uint64_t* cursor);
/* per-stage cost tables of the current stage, indexed by |a-b| (256 each) */
void orc_job_stage_luts(orc_job* j, float* my_cost256, float* guide_cost256);

/* Exact k-nearest resolved neighbours (declared canonical order: squared
 * distance, then dy, then dx, all ascending) of each query pixel against a
 * 1-byte-per-pixel resolved mask, non-tiling. Returns counts. */
void orc_knn(const uint8_t* mask, uint32_t w, uint32_t h, uint32_t n_q, const uint32_t* q_flat,
             uint32_t k, int32_t* nbrs, uint32_t* counts);

#ifdef __cplusplus
}
#endif
#endif
