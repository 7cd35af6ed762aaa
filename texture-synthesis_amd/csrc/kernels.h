// Kernel argument blocks and launchers of the patch-search engine.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace tsg {

constexpr int kMaxK = 64;          // nearest_neighbors supported on the device
constexpr uint32_t kBruteMax = 2048;  // resolved count up to which kNN scans the list
constexpr int kTableRadius = 64;   // offset table radius of the ring scan

// One Jacobi batch: items [item0, item1) of the current stage.
struct BatchArgs {
    uint32_t W, H, tiles_x, tiles_y;
    const uint64_t* mask;    // resolved bits, one u64 per 8x8 tile
    const uint4* state;      // per output pixel {colour, src xy, map ids, patch}
    const uint32_t* tg;      // target guide at the stage level (null = unguided)
    const MapDesc* maps;     // filtered example maps at the stage level
    uint32_t n_maps;
    const uint32_t* work;    // stage work list: flat pixel of item i
    const uint32_t* rlist;   // resolved list (first n_res entries committed)
    uint32_t n_res;
    const uint2* rcands;     // random candidates, item-major, m per item
    uint32_t rc_base;        // item index of rcands[0]
    uint32_t m;
    const int32_t* offtab;   // ring offsets (dx & 0xffff | dy << 16), canonical order
    uint32_t offtab_n;
    const float* lut;        // [0,256): colour cost by |a-b|; [256,512): guide cost
    uint32_t item0, item1;
    uint64_t stage_seed;
    uint32_t k;
    uint4* result;           // per batch item {src xy, map ids | fallback, patch, score bits}
    float* item_score;       // per stage item (stats)
    // test hooks: host-supplied neighbour / candidate lists (null in production)
    const int32_t* ext_nbrs;      // [item][k] packed dx | dy << 16
    const uint32_t* ext_nbr_count;
    const uint4* ext_cands;       // [item][cmax] {xy, map | kind << 31, patch, 0}
    const uint32_t* ext_cand_count;
    uint32_t ext_cmax;
    float* ext_full_cost;         // [item][cmax] full sums (no early exit)
    int32_t* dbg_nbrs;            // [item][k] kNN result dump (test hook)
    uint32_t* dbg_nbr_count;
    uint32_t* dbg_best;           // [item] winning candidate index (test hook)
    uint32_t knn_only;            // stop after the kNN dump (test hook)
    uint32_t exact;               // force the unpruned lane-per-candidate evaluation
    uint32_t lut_nan;             // the stage's cost tables hold a NaN (cauchy 0): exact path
    // commit fused into the resolve launch (last workgroup) / small-batch kernel
    uint32_t fused_commit;
    uint32_t first_new;           // items >= first_new are new pixels
    uint32_t* done_counter;       // zero between launches
    uint4* state_w;
    uint64_t* mask_w;
    float* score_w;
    unsigned long long* dbg_clk;  // [stage item][16] phase timestamps (profiling)
    unsigned long long* dbg_launch;  // {earliest workgroup start, latest end} of this launch (profiling)
    uint32_t team;                // waves per item (1, 2, 4, 8; 0 = 1)
    // neighbour lists computed ahead for stage items [pre_base, pre_end)
    const int32_t* pre_nbrs;      // [item - pre_base][k] packed dx | dy << 16
    const uint32_t* pre_cnt;
    uint32_t pre_base, pre_end;
    // tiling mode (null gmask: off): per 8x8 tile, the resolved pixels that
    // carry ghost copies one image width / height away, on the far side, while
    // within the edge margins x < xl, x > xr, y < yb, y > yt; neighbour records
    // and the own pattern wrap around the image
    uint64_t* gmask;
    int32_t xl, xr, yb, yt;
    // deferred commit (k_resolve; null state2: off). Items write their record
    // to state2 and `serial` to stamp; new pixels also set their bit in
    // pend_w. Readers take a neighbour from state2 when its stamp is
    // serial - 1 (the previous batch, which this launch commits) and search
    // the mask together with pend_r (its new pixels). Each workgroup then
    // commits a slice of the previous batch [prev0, prev1) into state_w and
    // mask_w, and clears pend_c for the batch before it [pp0, pp1).
    uint4* state2;
    uint32_t* stamp;
    uint32_t serial;
    const uint64_t* pend_r;
    uint64_t* pend_w;
    uint64_t* pend_c;
    uint32_t prev0, prev1, pp0, pp1;
};

// The commit a deferred batch left (its launch was the last of a series):
// items [prev0, prev1) from state2 into state and mask; pend_prev is cleared
// for them and pend_pp for items [pp0, pp1).
struct DeferFlushArgs {
    uint32_t W, tiles_x;
    const uint32_t* work;
    const uint4* state2;
    uint4* state;
    uint64_t* mask;
    uint64_t* pend_prev;
    uint64_t* pend_pp;
    uint32_t first_new;
    uint32_t prev0, prev1, pp0, pp1;
};

// kNN of stage items [item0, item1) ahead of their batches: item i sees the
// resolved set at the start of its batch, i.e. every pixel of `mask` (one
// u64 per 8x8 tile) and the pixels p of `mask_new` with rank[p] <
// n_at[i - item0] (rank = position in the resolved list). tile_min: least
// rank of each tile's mask_new pixels. mask_new may be null.
struct PrefetchArgs {
    uint32_t W, H, tiles_x, tiles_y;
    const uint64_t* mask;
    const uint64_t* mask_new;
    const uint32_t* tile_min;
    const uint32_t* rank;
    const uint32_t* rlist;
    const uint32_t* work;
    const uint32_t* n_at;         // [item - item0], or (null) from the batch table:
    const uint32_t* bounds;       // batch b holds items [bounds[b], bounds[b + 1]) (nb + 1 entries)
    const uint32_t* bnres;        // and starts with bnres[b] resolved pixels
    uint32_t nb;
    // items below redo_end redo resolved pixels: their dependency bits mark
    // neighbours redone by an earlier batch (rank - locked below the item's
    // batch start); the other items' bits mark pixels of mask_new
    uint32_t redo_end, locked;
    const int32_t* offtab;
    uint32_t offtab_n;
    uint32_t k;
    uint32_t item0, item1;
    int32_t* out_nbrs;            // [item - item0][k]
    uint32_t* out_cnt;
    // optional: per item, bit j set when neighbour j is a pixel of mask_new
    uint64_t* out_dep;            // [item - item0]
};

// A run of consecutive items [a.item0, a.item0 + n) of one segment (new
// pixels, or redo items with state_old) in one launch of persistent
// workgroups. Every item has its neighbour list
// (a.pre_nbrs) and a dependency mask dep[item - a.pre_base]: bit j set when
// neighbour j is a pixel committed by an earlier item of the run. Workgroups
// take items in order from a ticket counter, wait until each dependency's
// flag holds `epoch`, resolve, commit their item's pixel and publish it by
// writing `epoch` into its flag. Items wait only on earlier items, which
// running workgroups hold, so the run always drains. A wait longer than
// spin_limit polls sets bit 2 of *overflow and ends every wait of the launch
// (the results are then invalid and the host reports the failure).
struct FlowArgs {
    uint32_t* ticket;
    uint32_t ticket_base;         // counter value at the launch start
    uint32_t n;
    const uint64_t* dep;
    uint32_t* flag;               // per pixel: epoch of the run that committed it
    uint32_t epoch;
    uint32_t spin_limit;
    uint32_t* overflow;
    // a run of redo items: the state as the segment opened; neighbours
    // without a dependency bit are read there (the run rewrites the live
    // state), and commits touch neither mask nor score. Null: new pixels.
    const uint4* state_old;
};

struct RandArgs {
    const MapDesc* maps;
    uint32_t n_maps;
    uint64_t stage_seed;
    uint32_t item0, item1;
    uint32_t m;
    uint2* out;              // [item - item0][m]
    uint32_t* overflow;      // set when a rejection loop hits its bound
};

struct CommitArgs {
    uint32_t W, tiles_x;
    const MapDesc* maps;
    const uint32_t* work;
    const uint4* result;
    uint32_t item0, item1;
    uint32_t first_new;      // items >= first_new are new pixels
    uint4* state;
    uint64_t* mask;
    float* score;            // per pixel score of the first resolve
    uint64_t* gmask;         // tiling mode: ghost flags of new pixels (null: off)
};

struct RecolorArgs {
    const MapDesc* maps;
    const uint32_t* rlist;
    uint32_t n;
    uint4* state;
};

// Result maps of a finished job from its packed state: rgba (u32 per pixel),
// coord (x, y, map), ids (patch, map); resolved list order and the score of
// each listed pixel. Null outputs are skipped.
struct UnpackArgs {
    const uint4* state;
    uint32_t n;
    uint32_t* rgba;
    uint32_t* coord;
    uint32_t* ids;
    const uint32_t* rlist;
    const float* score;
    uint32_t n_res;
    float* rscore;
};
hipError_t launch_unpack(const UnpackArgs& a, hipStream_t s);

// Debug maps of a finished job (GeneratedImage::save_debug, ms.rs:603-653):
// patch and map ids coloured by seeded draws, and the uncertainty of each
// resolved pixel (out_unc is all zero for the others). Outputs are RGBA8
// packed in u32; null outputs are skipped.
hipError_t launch_id_maps(const uint4* state, uint32_t n, uint32_t* patch_rgba, uint32_t* map_rgba, hipStream_t s);
hipError_t launch_uncertainty(const uint32_t* rlist, const float* score, uint32_t n_res, uint32_t* out_unc,
                              hipStream_t s);
// CoordinateTransform::apply (lib.rs:178-210): out[i] = source map's pixel at
// (x, y) of triple i. A coordinate outside its map sets *bad.
struct XformSource {
    const uint32_t* px;
    uint32_t w, h;
};
hipError_t launch_xform_apply(const uint32_t* coord, uint32_t n, const XformSource* srcs, uint32_t n_src,
                              uint32_t* out, uint32_t* bad, hipStream_t s);

// per-stage statistics: sum and count of non-negative item scores
struct StatsArgs {
    const float* item_score;
    uint32_t n;
    double* sum;
    unsigned long long* count;
};

hipError_t launch_knn_prefetch(const PrefetchArgs& a, hipStream_t s);

// Stage statistics in a fixed order: chunks of kStatsChunk items, each summed
// by 256 threads (thread t takes items t, t + 256, ... in turn, then a tree
// over the threads), then the chunks' partial sums in chunk order. The
// single-job path and the work-queue engine use the same arithmetic, so a
// job's statistics do not depend on which ran it.
constexpr uint32_t kStatsChunk = 4096;
hipError_t launch_stats_chunks(const float* item_score, uint32_t n, double* part, unsigned long long* part_n,
                               double* sum, unsigned long long* count, hipStream_t s);

// ---- work-queue engine: many jobs in one launch of persistent workgroups ----
// Each job brings a program: a list of phases run in order, each a parallel
// loop over `units` work units (a phase starts once every unit of the one
// before it has finished). Workgroups take units from the jobs' cursors; the
// workgroup that finishes a phase's last unit does the phase's closing work
// and opens the next phase. A job runs on one XCD, so its data stays in
// that XCD's L2 from phase to phase.
enum QKind : uint32_t {
    Q_RESOLVE = 0,   // batch items [i0, i1), 8 / team items per unit
    Q_COMMIT = 1,    // batch items [i0, i1), kQElems per unit
    Q_RAND = 2,      // random candidates of items [i0, i1) into the job's chunk (rc_base = i0)
    Q_RECOLOR = 3,   // resolved-list positions [i0, i1) recoloured from the stage level
    Q_STATS = 4,     // stage items [i0, i1), kStatsChunk per unit; the closing workgroup sums the chunks
    Q_CLEAR = 5,     // tiles [i0, i1) of the prefetch window masks
    Q_MASKOR = 6,    // resolved-list positions [i0, i1) into the window masks
    Q_PREFETCH = 7,  // neighbour lists of items [i0, i1), 8 per unit
};
constexpr uint32_t kQElems = 512;

struct QPhase {
    uint32_t kind;
    uint32_t stage;    // stage of the job (keys the cost tables cached in LDS)
    uint32_t i0, i1;
    uint32_t n_res;    // resolve: resolved pixels at the batch start
    uint32_t team;     // resolve: waves per item
    uint32_t rc_base;  // resolve: item of the candidate chunk's first entry
    uint32_t pre_base, pre_end;  // resolve: items with neighbour lists (pre_end 0: none);
                                 // prefetch: pre_base = offset of the window's n_at, pre_end = 1
                                 // for a new-pixel window (mask_new, tile_min, rank)
    uint32_t units;
    uint32_t fuse;     // resolve: the closing workgroup commits the batch (no commit phase)
    uint32_t rep;      // resolve: items each team takes in turn within one unit
    uint32_t arg;      // index of the stage's batch arguments in the launch's argument array; a
                       // resolve phase's item range, n_res, team, rc_base and prefetch window
                       // (with the job's neighbour-list buffers) replace the stage's
};

struct QJob {
    const BatchArgs* stages;   // per stage slot
    const QPhase* phases;      // this launch's program
    uint32_t n_phases;
    uint32_t xcc;
    uint2* rc;
    uint4* result;
    uint32_t* overflow;
    uint64_t* mend;
    uint32_t* tmin;
    const uint32_t* n_at;
    int32_t* pre_nbrs;
    uint32_t* pre_cnt;
    const uint32_t* rank;
    double* st_part;
    unsigned long long* st_part_n;
    double* st_sum;            // per stage slot
    unsigned long long* st_cnt;
};

// Per job, per launch: cursor = phase << 32 | units handed out; finished =
// units of the current phase done. Zero at the launch start. Only the job's
// XCD writes them; each job's pair has a 256-byte block of its own, so no
// cache line is written from two XCDs (their L2s are not coherent).
struct QDyn {
    unsigned long long cursor;
    uint32_t finished;
    uint32_t pad[61];
};
static_assert(sizeof(QDyn) == 256, "one 256-byte block per job");

// Per XCD, in a 256-byte block of its own: jobs whose program is not through,
// units finished (idle workgroups give up, setting `error`, after idle_ticks
// of the 100 MHz clock without progress on their XCD).
struct QXcd {
    uint32_t active;
    uint32_t progress;
    uint32_t error;
    uint32_t units_resolve;          // units done, by kind (resolve, other)
    uint32_t units_other;
    uint32_t pad0;
    // 100 MHz clock ticks summed over the XCD's workgroups: searching for a
    // unit (including idle waits), resolve units, other units, phase closing
    unsigned long long t_find, t_resolve, t_other, t_close;
    uint32_t pad[50];
};
static_assert(sizeof(QXcd) == 256, "one 256-byte block per XCD");

// Per launch: jobs of XCD x are [job0[x], job0[x + 1]) (read only).
struct QCtl {
    uint32_t job0[9];
    uint32_t pad[55];
    QXcd x[8];
};
static_assert(sizeof(QCtl) == 256 * 9, "job table offsets, then one block per XCD");

// The launch's batch arguments: every resolve phase has its own entry (its
// batch, candidate chunk and neighbour-list window filled in), every other
// phase shares its stage's entry.
// A spread launch: every workgroup, on any XCD, may take units of any of its
// jobs (all listed as on XCD 0, whose control words every workgroup uses),
// and a unit's results are released to device scope (the L2 written back).
// For launches with fewer jobs than XCDs, which home placement would leave
// on a few XCDs.
constexpr uint32_t kEngineSpread = 1u;
// The engine's first evaluation chunk while a pruning bound exists, as a
// shift: 8 neighbours. With many jobs in flight the VALU binds, and pruning
// from the first 8 terms saves more than the extra gather round trip costs.
constexpr int kEngineFirstChunk = 3;
// The engine's offset-table kNN: offsets per lane per round trip (64 per
// round). A dense mask holds k neighbours within the first 64 or so, and the
// VALU binds, so rounds are kept short.
constexpr int kEngineKnnU = 1;

struct QEngineArgs {
    const QJob* jobs;
    QDyn* dyn;
    QCtl* ctl;
    const int32_t* offtab;
    uint32_t offtab_n;
    uint32_t flags;          // kEngineSpread: the launch's jobs are shared by every XCD
    uint64_t idle_ticks;
};

// Workgroups of the engine launch resident on the device at once.
uint32_t engine_max_wgs(bool guided);
hipError_t launch_engine(const QEngineArgs& e, const BatchArgs* args, uint32_t wgs, bool guided, hipStream_t s);
// (float)exp(-x_i) for the sweep arguments x_i, i in [i0, i0 + n) (test hook).
hipError_t launch_exp_sweep(uint64_t i0, uint64_t n, float* out, hipStream_t s);
// XCD of each of n workgroups of a probe launch (test hook: the engine's
// job placement relies on the hardware XCC id).
hipError_t launch_xcc_probe(uint32_t* out, uint32_t n, hipStream_t s);
// rank[rlist[q]] = q for q in [lo, hi)
hipError_t launch_rank(const uint32_t* rlist, uint32_t lo, uint32_t hi, uint32_t* rank, hipStream_t s);
// sets the mask bits of rlist[lo, hi); with tile_min, also the least q per tile
hipError_t launch_mask_or(const uint32_t* rlist, uint32_t lo, uint32_t hi, uint32_t W, uint32_t tiles_x,
                          uint64_t* mask, uint32_t* tile_min, hipStream_t s);

// Events the next resolve launch (launch_resolve / _small / _coop) records
// around its dispatch, for per-kernel profiling; null for none.
void set_launch_events(hipEvent_t start, hipEvent_t stop);

hipError_t configure_kernels();
size_t resolve_lds_bytes(bool guided);
size_t resolve_small_lds_bytes(bool guided);
uint32_t small_batch_max();
// items per workgroup of k_resolve (and per engine resolve unit and item turn) with `team` waves per item
uint32_t items_per_wg(uint32_t team);
// waves per item for a batch of n items (1, 2, 4 or 8)
uint32_t team_for(uint64_t n, uint32_t team_max, uint32_t waves);
// consecutive batches of at most small_batch_max() items, one workgroup
// Consecutive batches (each at most wgs * 16 items) in one cooperative launch
// of wgs workgroups of 16 waves; wgs <= coop_max_wgs(guided).
uint32_t coop_max_wgs(bool guided);
hipError_t launch_resolve_coop(const BatchArgs& a, const uint32_t* bounds, const uint32_t* nres, uint32_t n_batches,
                               uint32_t wgs, uint32_t team_max, uint32_t* overflow, bool guided, hipStream_t s);
hipError_t launch_resolve_small(const BatchArgs& a, const uint32_t* bounds, const uint32_t* nres,
                                uint32_t n_batches, bool guided, hipStream_t s);
hipError_t launch_resolve(const BatchArgs& a, bool guided, hipStream_t s);
// A dependency-ordered run (FlowArgs): wgs persistent workgroups of `team`
// waves; wgs <= flow_max_wgs(team, guided). Consumes n + wgs tickets.
uint32_t flow_max_wgs(uint32_t team, bool guided);
hipError_t launch_resolve_flow(const BatchArgs& a, const FlowArgs& f, uint32_t wgs, uint32_t team, bool guided,
                               hipStream_t s);
// Tile bitmask (one u64 per 8x8 tile) of a one-byte-per-pixel mask.
hipError_t launch_mask_tiles(const uint8_t* mask, uint32_t W, uint32_t H, uint32_t tiles_x, uint32_t tiles_y,
                             uint64_t* tiles, hipStream_t s);
// Packed kNN offsets (dx & 0xffff | dy << 16) to (x, y) int32 pairs, zero
// past each query's count.
hipError_t launch_knn_unpack(const uint32_t* q_flat, const int32_t* packed, const uint32_t* counts, uint32_t n_q,
                             uint32_t k, uint32_t W, int32_t* pairs, hipStream_t s);
hipError_t launch_stats(const StatsArgs& a, hipStream_t s);
hipError_t launch_rand(const RandArgs& a, hipStream_t s);
hipError_t launch_commit(const CommitArgs& a, hipStream_t s);
// Single-wave batches on persistent workgroups; persist_max_wgs: resident count.
uint32_t persist_max_wgs(bool guided);
hipError_t launch_resolve_persist(const BatchArgs& a, uint32_t wgs, bool guided, hipStream_t s);
hipError_t launch_defer_flush(const DeferFlushArgs& a, hipStream_t s);
hipError_t launch_state_snapshot(const uint32_t* rlist, uint32_t n, const uint4* state, uint4* old, hipStream_t s);
hipError_t launch_recolor(const RecolorArgs& a, hipStream_t s);
// separable resampling pass: vertical (u8 -> f32) and horizontal (f32 -> u8)
hipError_t launch_resample_v(const uint32_t* src, uint32_t w, uint32_t h, float4* dst, uint32_t nh,
                             const uint32_t* left, const uint32_t* cnt, const float* ws, uint32_t stride,
                             hipStream_t s);
hipError_t launch_resample_h(const float4* src, uint32_t w, uint32_t h, uint32_t* dst, uint32_t nw,
                             const uint32_t* left, const uint32_t* cnt, const float* ws, uint32_t stride,
                             hipStream_t s);

}  // namespace tsg
