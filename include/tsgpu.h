/*
 * tsgpu — C ABI of the MI355X (gfx950) patch-search engine for
 * example-based texture synthesis.
 *
 * The boundary replaces, inside EmbarkStudios/texture-synthesis (lib/src),
 *   Generator::resolve_random_batch        ms.rs:427-445
 *   Generator::resolve                     ms.rs:702-1051
 *   ImagePyramid::new / build_gaussian     img_pyramid.rs:7-37
 * as called by Session::run (session.rs:37-66) and SessionBuilder::build
 * (session.rs:336-448). All entry points take plain pointers and sizes;
 * images are row-major RGBA8 with stride 4*w (image::RgbaImage layout).
 * The caller owns every host buffer; buffers are copied during the call.
 * The library owns all device memory; handles are opaque.
 *
 * Status codes: 0 on success, negative on failure; tsg_last_error() returns
 * a thread-local message for the last failure on the calling thread.
 */
#ifndef TSGPU_H
#define TSGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes. -1..-5 mirror texture_synthesis::Error (errors.rs:38-57). */
enum tsg_status {
    TSG_OK = 0,
    TSG_ERR_INVALID_RANGE = -1,          /* Error::InvalidRange */
    TSG_ERR_NO_EXAMPLES = -2,            /* Error::NoExamples */
    TSG_ERR_EXAMPLE_GUIDE_MISMATCH = -3, /* Error::ExampleGuideMismatch */
    TSG_ERR_SIZE_MISMATCH = -4,          /* Error::SizeMismatch */
    TSG_ERR_MAPS_COUNT_MISMATCH = -5,    /* Error::MapsCountMismatch */
    TSG_ERR_INVALID_ARGUMENT = -10,      /* null pointer, zero size, bad index */
    TSG_ERR_UNSUPPORTED = -11,           /* outside the device path's limits */
    TSG_ERR_DEVICE = -20,                /* HIP runtime failure */
    TSG_ERR_OUT_OF_MEMORY = -21,         /* device allocation failed */
    TSG_ERR_SAMPLING = -22,              /* a sample mask admits no pixel */
    TSG_ERR_CANCELLED = -23,             /* the job was cancelled before it finished (tsg_batch_cancel) */
    TSG_ERR_ADMISSION_REJECTED = -24     /* the job can never fit the caller's device-memory budget */
};

/* GeneratorParams (ms.rs:17-42) lowered from Parameters (lib.rs:327-375).
 * tsg_params_default() gives the defaults of lib.rs:342-359. */
typedef struct tsg_params {
    uint32_t nearest_neighbors;       /* k, 1..64 on the device. default 50 */
    uint64_t random_sample_locations; /* m >= 1. default 50 */
    float cauchy_dispersion;          /* [0,1]. default 1.0 */
    float backtrack_percent;          /* p in [0,1]. default 0.5 */
    uint32_t backtrack_stages;        /* P. default 5 */
    float guide_alpha;                /* [0,1]. default 0.8 */
    uint64_t seed;                    /* default 0 */
    uint32_t tiling_mode;             /* 0/1. default 0. 1: ghost points at the far edge within 5%
                                       * margins (ms.rs:296-331), wrapped own pattern */
    uint64_t random_init;             /* UINT64_MAX = none (random_resolve None) */
    uint32_t max_thread_count;        /* accepted for API parity (>= 1); the device ignores it */
} tsg_params;

/* Jacobi batch schedule of the device path. All items of a batch read the
 * state as it stood at the batch start. px_per_batch > 0 fixes the batch
 * size (1 reproduces the reference at max_thread_count(1)); otherwise a
 * redo batch holds max(1, (resolved - locked) >> redo_shift) items and a
 * new-pixel batch max(1, (resolved - locked) >> new_shift), both capped at
 * max_batch when nonzero. */
typedef struct tsg_schedule {
    uint32_t px_per_batch;
    uint32_t redo_shift;
    uint32_t new_shift;
    uint32_t max_batch;
    uint32_t first_new_shift;  /* new-pixel shift of the first stage (0: new_shift) */
} tsg_schedule;

/* Launch structure of the device path. Results never depend on it (every
 * choice resolves the same batches against the same frozen states); it
 * trades launch count, latency and occupancy, and is what the autotuner
 * picks. tsg_tuning_default gives: small_max 4, fuse_max 0, team_max 8,
 * team_waves 4096, prefetch_max 1024, ahead 1, coop 0/0, flow 1, flow_wgs 0,
 * flow_team 4, defer 0, persist 0, flow_head 1024, flow_head_team 8.
 *   small_max    batches of at most this many items run in the one-workgroup
 *                loop kernel (0..16; 0 = never)
 *   fuse_max     larger batches of at most this many items commit in their
 *                resolve launch's last workgroup (0 = separate commit launch)
 *   team_max     waves per item of latency-bound batches (1, 2, 4 or 8)
 *   team_waves   resident waves the team size is sized against
 *   prefetch_max batches of at most this many items get their neighbour
 *                lists computed ahead, window by window (0 = never)
 *   ahead        1: a single-job run prepares each stage's content-independent
 *                work (random candidates, pick ranks, the neighbour lists of
 *                its leading latency-bound batches) during the previous
 *                stage on a second stream, in two buffer sets used in turn,
 *                when they fit in device memory; 2: one buffer set per stage,
 *                every item's neighbour list, prepared as far ahead as the
 *                host has reached (when the stages fit, else as 1); 0: per
 *                window on the job stream
 *   coop_wgs     workgroups (16 waves each) of the cooperative launch that
 *                runs consecutive latency-bound batches with grid barriers
 *                instead of two launches per batch (0 = off)
 *   coop_max     batches of at most this many items (<= 16 * coop_wgs) go
 *                to that launch
 *   flow         1: with `ahead`, each stage's leading run of latency-bound
 *                new-pixel batches goes to one launch of persistent
 *                workgroups that resolve the items in order, each waiting
 *                only for the neighbours committed inside the run (per-pixel
 *                flags); 2: as 1, and the leading run of latency-bound redo
 *                batches too, reading neighbours not redone before them from
 *                a snapshot of the state taken as the stage opens; 0: one
 *                launch (or loop step) per batch
 *   flow_wgs     workgroups of that launch (0: as many as are resident)
 *   flow_team    waves per item in that launch (1, 2, 4 or 8)
 *   defer        1: in a single-job run, a throughput-bound batch's commit is
 *                done by the next such batch's launch (its items read the
 *                previous batch's records from a second state array), so a
 *                series of batches needs one launch each instead of two;
 *                0 (default): a commit launch after every batch
 *   persist      1: a batch of single-wave items runs on resident
 *                workgroups, each over an equal slice whose items its waves
 *                take in turn (no wave waits for the slowest item of its
 *                workgroup); 0: one workgroup per 8 items (three resident
 *                per CU, where the persistent kernel's registers allow two)
 *   flow_head    leading items of each dependency-ordered run that go to a
 *                launch of their own, with teams of flow_head_team waves
 *                (the run's first items form its longest serial chain)
 *   flow_head_team  waves per item in that launch (1, 2, 4 or 8) */
typedef struct tsg_tuning {
    uint32_t small_max;
    uint32_t fuse_max;
    uint32_t team_max;
    uint32_t team_waves;
    uint32_t prefetch_max;
    uint32_t ahead;
    uint32_t coop_wgs;
    uint32_t coop_max;
    uint32_t flow;
    uint32_t flow_wgs;
    uint32_t flow_team;
    uint32_t defer;
    uint32_t persist;
    uint32_t flow_head;
    uint32_t flow_head_team;
} tsg_tuning;

typedef struct tsg_image {
    uint32_t w, h;
    const uint8_t* rgba; /* w*h*4 bytes */
} tsg_image;

/* SamplingMethod (lib.rs:458-470) */
enum tsg_sample_kind { TSG_SAMPLE_ALL = 0, TSG_SAMPLE_IGNORE = 1, TSG_SAMPLE_IMAGE = 2 };

typedef struct tsg_ctx tsg_ctx;
typedef struct tsg_job tsg_job;

/* Progress, as GeneratorProgress::update (session.rs:527-558): called on the
 * caller's thread between batches whenever the rounded overall percentage
 * changes. `rgba` (out_w*out_h*4) is valid only during the call. */
typedef void (*tsg_progress_fn)(void* user, const uint8_t* rgba, uint32_t w, uint32_t h, uint64_t total_current,
                                uint64_t total, uint64_t stage_current, uint64_t stage_total);

void tsg_params_default(tsg_params* p);
void tsg_schedule_default(tsg_schedule* s);
void tsg_tuning_default(tsg_tuning* t);
const char* tsg_last_error(void);
const char* tsg_version(void);

/* Context: one HIP device and stream. */
int tsg_ctx_create(int device, tsg_ctx** out);
void tsg_ctx_destroy(tsg_ctx* ctx);

/* SessionBuilder::build (session.rs:336-448) for already-decoded images.
 * examples: finest example images at their final size, in add order
 *           (load_image / resize_input are the caller's, utils.rs:55-80).
 * sample_kind[i] / sample_masks[i]: per example; sample_kind NULL = all ALL.
 *           A TSG_SAMPLE_IMAGE mask has the example's size.
 * example_guides: NULL, or one guide per example (all-or-none,
 *           session.rs:514-521); requires target_guide.
 * target_guide: NULL, or an out_w x out_h guide.
 * inpaint_mask: NULL, or an out_w x out_h mask; pixels with R < 255 are
 *           synthesised (ms.rs:271-279); examples[inpaint_index] is then the
 *           initial colour map and must be out_w x out_h.
 * Validation errors are returned before any device work. */
int tsg_job_create(tsg_ctx* ctx, const tsg_params* params, uint32_t n_examples, const tsg_image* examples,
                   const int* sample_kind, const tsg_image* sample_masks, const tsg_image* example_guides,
                   const tsg_image* target_guide, const tsg_image* inpaint_mask, uint32_t inpaint_index,
                   uint32_t out_w, uint32_t out_h, tsg_job** out);
/* Overrides the default Jacobi schedule (before tsg_job_run). */
int tsg_job_set_schedule(tsg_job* job, const tsg_schedule* s);
/* Overrides the launch structure (before tsg_job_run). A new job starts
 * with the one its context stored for its shape, else tsg_tuning_default. */
int tsg_job_set_tuning(tsg_job* job, const tsg_tuning* t);
int tsg_job_get_tuning(const tsg_job* job, tsg_tuning* out);
/* Autotuner cache (no counterpart in the reference): stores t for jobs of
 * the same shape as `like` (output size, k, m, sampled maps, guided, tiling,
 * stages) created on ctx from now on; the search that picks t is the
 * caller's (tsgpu.autotune times candidate structures on the device). */
int tsg_ctx_store_tuning(tsg_ctx* ctx, const tsg_job* like, const tsg_tuning* t);
int tsg_ctx_clear_tunings(tsg_ctx* ctx);
/* Session::run (session.rs:37-66): random init, then every stage. */
int tsg_job_run(tsg_job* job, tsg_progress_fn progress, void* user);
/* GeneratedImage accessors (lib.rs:378-455, ms.rs:605-684). Any pointer may
 * be NULL. rgba: w*h*4. coord: w*h*3 u32 (x, y, map) — the .xform body.
 * ids: w*h*2 u32 (patch id, map id). resolved_flat / resolved_score: the
 * resolved list in order (w*h entries). */
/* Compact form of tsg_job_download: the packed per-pixel state, four u32
 * per pixel {rgba (R in the low byte), x | y << 16, coord-map id |
 * id-map id << 16, patch id}, plus the resolved list and its scores (each
 * pointer may be NULL). */
int tsg_job_download_packed(tsg_job* job, uint32_t* state, uint32_t* resolved_flat, float* resolved_score);
/* GeneratedImage::save_debug's maps of a finished job (Generator::get_id_maps
 * and get_uncertainty_map, ms.rs:603-653), RGBA8 w*h*4 each, any may be NULL:
 * patch ids and map ids coloured by seeded gen_range(0..255) draws, and the
 * uncertainty [n, 255 - n, 0, 255] with n = (min(score, 1) * 255) as u8 of
 * every resolved pixel (others stay 0). */
int tsg_job_debug_maps(tsg_job* job, uint8_t* patch_ids, uint8_t* map_ids, uint8_t* uncertainty);
/* CoordinateTransform::apply (lib.rs:178-210) on the device: out (w*h*4
 * RGBA8) pixel i = sources[map][y][x] for coord triple i = (x, y, map), the
 * .xform body. Sources at the transform's original sizes (resizing them is
 * the caller's, as load_image does). A coordinate outside its source (the
 * reference panics) returns TSG_ERR_INVALID_ARGUMENT. */
int tsg_xform_apply(tsg_ctx* ctx, const uint32_t* coord, uint32_t w, uint32_t h, uint32_t n_sources,
                    const tsg_image* sources, uint8_t* out);
int tsg_job_download(tsg_job* job, uint8_t* rgba, uint32_t* coord, uint32_t* ids, uint32_t* resolved_flat,
                     float* resolved_score);
/* Per stage (index 0 = p_stage P): sum of the winning costs of items
 * resolved by pattern matching with a numeric cost (NaN weights, from a zero
 * mean neighbour distance, give NaN costs), their count, and items
 * processed. Returns the number of stages written (<= cap). */
int tsg_job_stage_stats(tsg_job* job, double* sums, uint64_t* counts, uint64_t* items, uint32_t cap);
/* Device time of the last tsg_job_run in milliseconds, and batch count. */
int tsg_job_timing(tsg_job* job, double* run_ms, uint64_t* n_batches);
/* Per-kernel device timing of the next tsg_job_run (HIP events on the job's
 * stream around every launch). Level -1 records only the resolve kernels'
 * execution spans on the device clock (no events; light enough for timed
 * runs: the other kernels then report nothing). Level L >= 2 also records, for each work item
 * of stage L - 2 (0 = the first), device timestamps (100 MHz wall clock) at
 * the phase boundaries of its resolve: start, kNN, neighbour state and
 * coherence candidates, first candidate group, all groups, end. */
int tsg_job_set_profiling(tsg_job* job, int level);
/* Phase record of the first n_items work items of the recorded stage (16 u64
 * each) after a level >= 2 profiled run: [0..6] the timestamps above (slot 3
 * unused), [7] kNN search done (in-kernel searches only), [8] ring-search
 * steps, [9] keys buffered before the final selection. */
int tsg_job_phase_clocks(tsg_job* job, uint64_t* out, uint64_t n_items);
/* After a profiled run, per kernel [resolve (one launch per batch), commit,
 * random candidates, recolour, stats, kNN prefetch, resolve runs (several
 * small batches per launch)]: summed milliseconds, items processed,
 * launches. Each array holds 7 entries. The resolve kernels' milliseconds
 * are their execution spans on the device clock (earliest workgroup start
 * to latest workgroup end); the others' come from events around them. */
int tsg_job_kernel_stats(tsg_job* job, double* ms, uint64_t* items, uint64_t* launches);
/* Frees the job. A job submitted to a tsg_batch belongs to the batch until
 * tsg_batch_take hands it back: destroying it before is refused (the job
 * stays, tsg_last_error says why) and the batch frees it with itself. */
void tsg_job_destroy(tsg_job* job);

/* ---- many syntheses at once (no counterpart in the reference: serving and
 *      batch callers that run many Session::run at the same time) ----
 * A tsg_batch runs its admitted jobs together. Each step (a "round") plans
 * every job's next phases exactly as a run alone would batch them (random
 * candidates, each Jacobi batch's resolve and commit, stage statistics) and
 * runs all of them in one launch of a persistent work-queue kernel, whose
 * workgroups claim the jobs' work units as they free up. A job's result is
 * identical to tsg_job_run on it alone.
 * Jobs are created with tsg_job_create and handed over with
 * tsg_batch_submit; once finished they are taken back with tsg_batch_take
 * (then downloaded and destroyed by the caller). */
typedef struct tsg_batch tsg_batch;
/* Device bytes held by a job (pyramids, state, scratch). */
int tsg_job_footprint(const tsg_job* job, uint64_t* bytes);
/* Device bytes a job of these dimensions holds while it runs under a
 * tsg_batch, before it is created (for admission by memory). example_wh:
 * n_examples (w, h) pairs; sample_kind may be NULL (all TSG_SAMPLE_ALL);
 * guide_wh: NULL when unguided, else the example guides' (w, h) pairs. */
int tsg_footprint_estimate(const tsg_params* p, uint32_t n_examples, const uint32_t* example_wh,
                           const int* sample_kind, const uint32_t* guide_wh, uint32_t out_w, uint32_t out_h,
                           uint64_t* bytes);
/* At most max_active jobs run at a time; the rest wait in submission order. */
int tsg_batch_create(tsg_ctx* ctx, uint32_t max_active, tsg_batch** out);
int tsg_batch_submit(tsg_batch* b, tsg_job* job, uint64_t* ticket);
/* Resolve items each job may plan per round (0, the default: everything its
 * pick orders allow). A bound keeps rounds short, so that jobs submitted
 * while others run start soon (serving); a job still plans at least one
 * Jacobi batch per round. Results do not depend on it. */
int tsg_batch_set_quantum(tsg_batch* b, uint64_t items);
/* A job that starts while fewer than `jobs` jobs are in the batch (running
 * and waiting, itself included) runs alone instead: to its end, on the
 * single-job path with the whole device (tsg_job_run's launch sequence),
 * before the round's engine launch; a round's alone jobs run smallest first.
 * Its stage-ahead buffers, beyond tsg_footprint_estimate, are used only when
 * they fit ahead_bytes, and are released when its run ends. jobs = 0, the
 * default: every job runs on the engine. Under light load a lone job's chain
 * of small batches is what bounds its latency, and the single-job path
 * shortens that chain; under load the engine's throughput wins. Results do
 * not depend on it. */
int tsg_batch_set_alone_below(tsg_batch* b, uint32_t jobs, uint64_t ahead_bytes);
/* Runs up to max_rounds rounds; in_flight = jobs not yet reported finished. */
int tsg_batch_step(tsg_batch* b, uint32_t max_rounds, uint32_t* in_flight);
/* Runs until every submitted job has finished. */
int tsg_batch_run_all(tsg_batch* b);
/* Finished jobs not yet taken: tickets and their run status (0 = ok). */
int tsg_batch_poll(tsg_batch* b, uint64_t* tickets, int* status, uint32_t cap, uint32_t* n);
int tsg_batch_take(tsg_batch* b, uint64_t ticket, tsg_job** job);
/* Rounds and engine launches so far, engine phases run, and jobs run alone
 * (tsg_batch_set_alone_below); any pointer may be NULL. */
int tsg_batch_stats(tsg_batch* b, uint64_t* rounds, uint64_t* launches, uint64_t* units, uint64_t* alone);
/* Cancels a job that has not finished: a waiting job never starts, a
 * running one stops between launches. Either is then reported finished by
 * tsg_batch_poll with status TSG_ERR_CANCELLED and taken back as usual (its
 * results are not usable). The other jobs are not affected. */
int tsg_batch_cancel(tsg_batch* b, uint64_t ticket);
/* Destroys the batch and every job it still holds. */
void tsg_batch_destroy(tsg_batch* b);

/* Resampling filters of image 0.23's imageops::resize used on the path:
 * Triangle (inpaint masks, ms.rs:242-263), CatmullRom (load_image,
 * utils.rs:55-80), Gaussian (ImagePyramid levels, img_pyramid.rs:27-32). */
enum tsg_filter { TSG_FILTER_TRIANGLE = 0, TSG_FILTER_CATMULLROM = 1, TSG_FILTER_GAUSSIAN = 2 };

/* load_image's resize step (utils.rs:55-80: image::imageops::resize with
 * the given filter, vertical pass to f32, horizontal pass clamped and
 * rounded to u8) on the device: out is nw*nh*4 RGBA8. An image already of
 * the asked size is copied unchanged, as load_image does (utils.rs:66). */
int tsg_resize(tsg_ctx* ctx, const tsg_image* img, uint32_t nw, uint32_t nh, int filter, uint8_t* out);

/* ---- test and tooling entry points (no counterpart in the reference API;
 *      they expose single steps of the path for parity tests) ---- */

/* Where the engine's time went, summed over the batch's launches: workgroup
 * milliseconds searching for units (idle waits included), in resolve units,
 * in other units, closing phases; the counts of resolve units and other
 * units; then host milliseconds admitting jobs, planning their programs,
 * packing and queueing the launch block, from the launches to their
 * completion, and finishing jobs (11 values). */
int tsg_batch_engine_profile(tsg_batch* b, double* out11);
/* Test hook of the work-queue engine: at most wgs_cap workgroups per launch
 * (0: as many as are resident); flags bit 0 spreads every launch over all
 * XCDs (by default only launches of fewer than 8 jobs are). */
int tsg_batch_debug_engine(tsg_batch* b, uint32_t wgs_cap, uint32_t flags);

/* Failure injection for tests of error handling: the library's next `count`
 * device allocations fail as if the device were out of memory (0 turns it
 * off). Process-wide. */
int tsg_debug_fail_allocs(int64_t count);

/* Device value of the neighbour weights' (float)exp(-1.0 * x) (ms.rs:1198-1203;
 * the library's own double-double exp) for the sweep arguments x_i, i in
 * [i0, i0 + n), n <= 2^28: t = ((i * 0x9E3779B97F4A7C15) mod 2^64 >> 11) *
 * 2^-53, x_i = 8 t when i % 4 != 0, else 8 + 56 t (f64, no contraction).
 * Test hook for comparing with the host C library's exp, as the reference
 * calls it. tsg_debug_exp_sweep_host computes the same on the host. */
int tsg_debug_exp_sweep(tsg_ctx* ctx, uint64_t i0, uint64_t n, float* out);
int tsg_debug_exp_sweep_host(uint64_t i0, uint64_t n, float* out);

/* ImagePyramid::new(img, Some(levels)) on the device: writes max(levels,1)
 * full-size RGBA8 levels, coarsest first (img_pyramid.rs:20-37). */
int tsg_pyramid_build(tsg_ctx* ctx, const tsg_image* img, uint32_t levels, uint8_t* out);

/* k nearest resolved neighbours (canonical order: d^2, dy, dx) of each
 * query pixel against a 1-byte-per-pixel mask (nonzero = resolved),
 * non-tiling. nbrs: n_q*k (x, y) int32 pairs; counts: n_q. Replaces
 * TreeGrid::get_k_nearest_neighbors (ms.rs:1363-1530). */
int tsg_knn_batch(tsg_ctx* ctx, const uint8_t* mask, uint32_t w, uint32_t h, uint32_t n_q, const uint32_t* q_flat,
                  uint32_t k, int32_t* nbrs, uint32_t* counts);

/* tsg_knn_batch without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, results and errors as tsg_knn_batch; every
 * device buffer (staged mask and queries, tile bits, resolved list, lists,
 * results) lives in a fixed 32 MiB device array per GPU, so the call
 * returns TSG_ERR_UNSUPPORTED when w*h + w*h/8 + n_q*(12*k + 24) + 8 KiB
 * (each part rounded up to 256 bytes) exceeds 32 MiB. All calls on one GPU
 * share that array: calls that overlap on one GPU race. */
int tsg_knn_batch_noalloc(tsg_ctx* ctx, const uint8_t* mask, uint32_t w, uint32_t h, uint32_t n_q,
                          const uint32_t* q_flat, uint32_t k, int32_t* nbrs, uint32_t* counts);

/* Random candidates (ms.rs:549-599) of stage items [item0, item0+n):
 * stream Pcg32::seed_from_u64(stage_seed + i + 1). maps / sample masks are
 * the filtered example maps. out: n*m (x, y, map) u32 triples. */
int tsg_random_candidates(tsg_ctx* ctx, uint32_t n_maps, const tsg_image* maps, const tsg_image* sample_masks,
                          uint64_t stage_seed, uint32_t item0, uint32_t n, uint32_t m, uint32_t* out);

/* Frozen-state scoring (find_best_match / better_match, ms.rs:1184-1288)
 * with host-supplied neighbour and candidate lists.
 * state: color (w*h*4 u8), coord (w*h*3 u32), ids (w*h*2 u32) of the output.
 * maps / guides: filtered example maps at the stage level (+ their guides,
 * or NULL); target_guide: w x h level image or NULL.
 * lut_color / lut_guide: 256-entry cost tables by |a-b| (the stage's
 * (1-alpha)*cauchy and alpha*l2; lut_guide NULL when unguided).
 * per item i: flat[i]; nbr_count[i] and nbrs[i*k + j] = (x, y);
 * cand_count[i] and cands[(i*cmax + c)*6] = (x, y, map, kind 0=coherence
 * 1=random, patch, patch map).
 * outputs: full_cost[i*cmax + c] (no early exit), best[i], best_score[i]
 * from the unpruned evaluation; best_pruned / best_score_pruned (nullable)
 * from the production evaluation (coherence dedup + exact pruning) with
 * `team` wavefronts per item (0 or 1, 2, 4, 8; the candidate list is split
 * between them).
 * Coherence candidates (kind 0, at most 64) must precede random ones. */
int tsg_score_batch(tsg_ctx* ctx, uint32_t w, uint32_t h, const uint8_t* color, const uint32_t* coord,
                    const uint32_t* ids, uint32_t n_maps, const tsg_image* maps, const tsg_image* guides,
                    const tsg_image* target_guide, const float* lut_color, const float* lut_guide, uint32_t k,
                    uint32_t cmax, uint32_t n, const uint32_t* flat, const uint32_t* nbr_count, const int32_t* nbrs,
                    const uint32_t* cand_count, const int32_t* cands, float* full_cost, uint32_t* best,
                    float* best_score, uint32_t* best_pruned, float* best_score_pruned, uint32_t team);

#ifdef __cplusplus
}
#endif
#endif /* TSGPU_H */
