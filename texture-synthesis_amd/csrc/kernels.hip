// gfx950 kernels of the patch-search engine.
//
// One wavefront resolves one output pixel (one "work item" of a stage):
//   kNN over the resolved-pixel mask            (replaces TreeGrid, ms.rs:1363-1530)
//   distance weights in f64                     (ms.rs:405-425, 1198-1203)
//   coherence candidates from neighbour state   (ms.rs:496-547)
//   + precomputed random candidates             (ms.rs:549-599, k_rand below)
//   Cauchy-cost compare, one lane per candidate (better_match, ms.rs:1227-1288)
//   argmin with the reference's replacement rule (find_best_match, ms.rs:1184-1224)
// A batch of items reads the state frozen at the batch start; k_commit then
// writes it (Generator::update, ms.rs:334-377).
//
// The cost sums keep the reference's f32 order: per neighbour four colour
// terms (then four guide terms) into one accumulator, then score += s * w as
// a separate multiply and add. Build with -ffp-contract=off.
#include <float.h>
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_ext.h>
#include <math.h>

#include "exp_dd.h"
#include "kernels.h"

namespace tsg {

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerWG = 8;
constexpr int kSmallWaves = 16;  // items per batch of k_resolve_small
constexpr int kKeyCap = 256;
// Interleaved copies of each 256-entry cost table in LDS (lane l reads copy
// l % copies): 8 for an unguided job, 4 of each of a guided job's two tables.
// Either way an 8-wave workgroup needs 51 KB, so three are resident per CU
// (24 waves). 32 copies fitted a single guided workgroup per CU.
template <bool GUIDED>
constexpr int lut_copies() {
    return GUIDED ? 4 : 8;
}
// LDS of the cost tables: a guided job's two tables (4 copies each) take what
// the unguided table (8 copies) takes.
__host__ __device__ constexpr size_t lut_bytes(bool) { return 256 * 8 * sizeof(float); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A copy of a record that every lane reads at the same address (job
// programs, phases, argument blocks), each 32-bit word moved to a scalar
// register as it arrives. After an acquire fence the compiler cannot treat
// such records as unchanged since the launch, so it reads them with vector
// loads; held as loaded, their fields took some 30 VGPRs plus spills through
// the whole resolve loop, while as scalars they cost no VGPRs.
template <class T>
__device__ __forceinline__ T ld_uniform(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "record of whole words");
    T v;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(p);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&v);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = (uint32_t)__builtin_amdgcn_readfirstlane((int)src[i]);
    return v;
}

// Loads and stores of device memory through the global address space. For a
// pointer read out of an argument record the compiler cannot tell device
// memory from LDS and emits flat instructions; a flat access also counts as
// an LDS access, so every later wait for an LDS read would wait for it too,
// which serialises a lane's gathers behind its LDS reads. Only for pointers
// to device buffers (never LDS).
template <class T>
struct same_type {
    using type = T;
};
template <class T>
__device__ __forceinline__ T gld(const T* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ void gst(T* p, typename same_type<T>::type v) {
    *(__attribute__((address_space(1))) T*)p = v;
}
// Element of a device buffer at a 32-bit byte offset: with a uniform base
// the load takes the scalar-base form, and the offset needs no 64-bit
// arithmetic. Offsets fit: job creation bounds the output at 16384^2 (per
// pixel tables of 4 bytes, tile masks of 2^22 words) and an example level at
// 2^30 pixels.
template <class T>
__device__ __forceinline__ T gld_at(const T* base, uint32_t byte_off) {
    return *(const __attribute__((address_space(1))) T*)(reinterpret_cast<const char*>(base) + byte_off);
}

// HIP's vector structs move as the native vectors they hold.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint2 gld(const uint2* p) {
    const u32x2 v = *(const __attribute__((address_space(1))) u32x2*)p;
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint4 gld(const uint4* p) {
    const u32x4 v = *(const __attribute__((address_space(1))) u32x4*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst(uint4* p, uint4 v) {
    u32x4 w;
    w.x = v.x;
    w.y = v.y;
    w.z = v.z;
    w.w = v.w;
    *(__attribute__((address_space(1))) u32x4*)p = w;
}

// Set bits of m below the calling lane (v_mbcnt_lo, v_mbcnt_hi): no 64-bit
// lane mask is held in registers.
__device__ __forceinline__ uint32_t bits_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Canonical neighbour order: squared distance, then dy, then dx, ascending.
__device__ __forceinline__ uint64_t make_key(int dx, int dy) {
    uint64_t d2 = (uint64_t)((int64_t)dx * dx + (int64_t)dy * dy);
    return (d2 << 32) | ((uint64_t)(uint32_t)(dy + 32768) << 16) | (uint64_t)(uint32_t)(dx + 32768);
}
__device__ __forceinline__ int32_t key_to_packed(uint64_t key) {
    int dx = (int)(key & 0xffffu) - 32768;
    int dy = (int)((key >> 16) & 0xffffu) - 32768;
    return (int32_t)(((uint32_t)dx & 0xffffu) | ((uint32_t)dy << 16));
}
__device__ __forceinline__ int packed_dx(int32_t o) { return (int)(int16_t)(o & 0xffff); }
__device__ __forceinline__ int packed_dy(int32_t o) { return o >> 16; }

constexpr uint32_t kNone = 0xffffffffu;
constexpr int kTermCap = 512;  // terms per chunk of the pruned evaluation
constexpr uint32_t kLead = 8;  // coherence candidates evaluated before the rest

struct WaveLds {
    union {
        uint64_t keys[kKeyCap];   // kNN selection
        float terms[kTermCap];    // per-chunk cost terms (candidate-major)
    };
    static_assert(kTermCap * 4 <= kKeyCap * 8, "terms share the key buffer");
    // the distances serve the weights only, which are done before the
    // evaluation's active candidates are written
    union {
        double dist[kMaxK];
        struct {                      // active candidates of the pruned evaluation
            uint32_t act_xy[kWave];
            uint32_t act_ms[kWave];   // map | 1 << 31 for mirrored (random) patterns
            uint32_t act_idx[kWave];  // index in the reference's candidate list
            float act_sum[kWave];
        };
    };
    int32_t nd[kMaxK];
    float w[kMaxK];
    uint32_t pat[kMaxK];
    uint32_t gpat[kMaxK];
    uint4 coh[kMaxK];
    double avg;
    // team exchange: neighbour count (leader), best (cost, index) per member,
    // and (leader's copy) the team's least complete (cost, index) so far,
    // packed as cost bits << 32 | index for a 64-bit minimum
    uint32_t t_kf;
    float t_B;
    uint32_t t_Bi;
    uint32_t t_xy, t_ms;  // the member's best candidate's coordinate and map word
    unsigned long long t_bound;
};

// ---- k-smallest selection over canonical keys, one wavefront ----
struct Sel {
    uint32_t count;
    uint64_t kth;  // current k-th smallest key once `count` reached k
};

__device__ __noinline__ void wave_sort(uint64_t* keys, int lane, int n) {
    // bitonic sort of keys[0, n), n a power of two in [2, kKeyCap]
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < n / 2; t += kWave) {
                int i = ((t & ~(stride - 1)) << 1) | (t & (stride - 1));
                int j = i + stride;
                uint64_t a = keys[i], b = keys[j];
                bool asc = (i & size) == 0;
                if ((a > b) == asc) {
                    keys[i] = b;
                    keys[j] = a;
                }
            }
            wave_sync();
        }
    }
}

// Bitonic sort of one u64 key per lane, ascending in lane order.
__device__ __forceinline__ uint64_t wave_sort64(uint64_t v, int lane) {
    for (int size = 2; size <= kWave; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const uint64_t o = ((uint64_t)(uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), stride) << 32) |
                               (uint32_t)__shfl_xor((int)(uint32_t)v, stride);
            const bool up = (lane & size) == 0 || size == kWave;
            const bool keep_min = ((lane & stride) == 0) == up;
            v = keep_min ? (o < v ? o : v) : (o > v ? o : v);
        }
    }
    return v;
}

// Keeps the k smallest buffered keys, sorted, in keys[0, min(count, k)).
// Up to 64 keys: one register sort. More: the k-th smallest squared
// distance T is found by bisection (ballot counts), the keys with d^2 <= T
// (k plus ties) are compacted and register-sorted; a tie group that pushes
// them past 64 takes the full LDS sort instead.
// The selection state is passed and returned by value: taken by reference it
// would live in scratch memory (its address escapes to this out-of-line call),
// and every push in the search loops would read and write it there.
__device__ __noinline__ Sel sel_compact(uint64_t* keys, Sel s, uint32_t k, int lane) {
    const uint32_t cnt = s.count;
    if (cnt <= (uint32_t)kWave) {
        uint64_t v = (uint32_t)lane < cnt ? keys[lane] : ~0ull;
        v = wave_sort64(v, lane);
        wave_sync();
        keys[lane] = v;
        wave_sync();
        s.count = cnt < k ? cnt : k;
        s.kth = s.count == k ? keys[k - 1] : ~0ull;
        return s;
    }
    uint64_t v[kKeyCap / kWave];
    uint32_t hi = 0;
#pragma unroll
    for (int r = 0; r < kKeyCap / kWave; ++r) {
        const uint32_t idx = (uint32_t)(lane + r * kWave);
        v[r] = idx < cnt ? keys[idx] : ~0ull;
        if (idx < cnt) hi = max(hi, (uint32_t)(v[r] >> 32));
    }
    for (int off = kWave / 2; off > 0; off >>= 1) hi = max(hi, (uint32_t)__shfl_xor((int)hi, off));
    uint32_t lo = 0;
    while (lo < hi) {  // smallest T with #{d^2 <= T} >= k
        const uint32_t mid = lo + ((hi - lo) >> 1);
        uint32_t c = 0;
#pragma unroll
        for (int r = 0; r < kKeyCap / kWave; ++r) c += (uint32_t)__popcll(__ballot((uint32_t)(v[r] >> 32) <= mid));
        if (c >= k) hi = mid;
        else lo = mid + 1;
    }
    uint32_t n_le = 0;
#pragma unroll
    for (int r = 0; r < kKeyCap / kWave; ++r) n_le += (uint32_t)__popcll(__ballot((uint32_t)(v[r] >> 32) <= lo));
    wave_sync();
    if (n_le <= (uint32_t)kWave) {
        uint32_t base = 0;
#pragma unroll
        for (int r = 0; r < kKeyCap / kWave; ++r) {
            const bool sel = (uint32_t)(v[r] >> 32) <= lo;
            const uint64_t m = __ballot(sel);
            if (sel) keys[base + bits_below(m)] = v[r];
            base += (uint32_t)__popcll(m);
        }
        wave_sync();
        uint64_t w = (uint32_t)lane < n_le ? keys[lane] : ~0ull;
        w = wave_sort64(w, lane);
        keys[lane] = w;
        wave_sync();
    } else {
        for (int t = lane; t < kKeyCap; t += kWave)
            if ((uint32_t)t >= cnt) keys[t] = ~0ull;
        wave_sync();
        wave_sort(keys, lane, kKeyCap);
    }
    s.count = k;
    s.kth = keys[k - 1];
    return s;
}

__device__ __forceinline__ void sel_push(uint64_t* keys, Sel& s, uint32_t k, bool valid, uint64_t key, int lane) {
    bool take = valid && key < s.kth;
    uint64_t m = __ballot(take);
    if (take) keys[s.count + bits_below(m)] = key;
    s.count += (uint32_t)__popcll(m);
    if (s.count > kKeyCap - kWave) {
        wave_sync();
        s = sel_compact(keys, s, k, lane);
    }
}

__device__ __forceinline__ bool px_set(const uint64_t* mask, uint32_t tiles_x, int X, int Y) {
    uint64_t wd = gld_at(mask, ((uint32_t)(Y >> 3) * tiles_x + (uint32_t)(X >> 3)) * 8u);
    return (wd >> (((Y & 7) << 3) | (X & 7))) & 1ull;
}

// What the kNN routes read: output size, resolved-pixel mask (one u64 per
// 8x8 tile), resolved list, offset table. With `mask_new`, a pixel set there
// also counts when rank[p] < n_at (resolved before an earlier moment);
// tile_min[t] is the least rank among tile t's mask_new pixels.
// In tiling mode (gmask set) the points also include the ghost copies of
// flush_resolved (ms.rs:296-331): a pixel with its gmask bit set and x < xl
// (x > xr) also stands at x + W (x - W); with y < yb (y > yt) at y + H
// (y - H); there are no diagonal copies.
struct KnnView {
    uint32_t W, H, tiles_x, tiles_y;
    const uint64_t* mask;
    const uint32_t* rlist;
    const int32_t* offtab;
    uint32_t offtab_n;
    const uint64_t* mask_new;
    const uint32_t* tile_min;
    const uint32_t* rank;
    uint32_t n_at;
    const uint64_t* gmask;
    int32_t xl, xr, yb, yt;
    const uint64_t* pend;  // more resolved bits (a deferred commit's new pixels), or null
    const int32_t* offc;   // the table's first kOffCache offsets in LDS, or null
};

constexpr uint32_t kOffCache = 256;  // one table round

// The workgroup's LDS copy of the offset table's first round (filled by
// setup_lds in every resolve kernel).
__device__ __forceinline__ int32_t* offtab_lds() {
    __shared__ int32_t cache[kOffCache];
    return cache;
}

// The resolved bits of tile ti.
__device__ __forceinline__ uint64_t mask_word(const KnnView& v, uint32_t ti) {
    uint64_t w = gld_at(v.mask, ti * 8u);
    if (v.pend) w |= gld_at(v.pend, ti * 8u);
    return w;
}

__device__ __forceinline__ bool view_set(const KnnView& v, int X, int Y) {
    return (mask_word(v, (uint32_t)(Y >> 3) * v.tiles_x + (uint32_t)(X >> 3)) >> (((Y & 7) << 3) | (X & 7))) & 1ull;
}

// Whether a point stands at (X, Y), ghosts included (no mask_new).
__device__ __forceinline__ bool point_at(const KnnView& v, int X, int Y) {
    const bool xin = X >= 0 && X < (int)v.W, yin = Y >= 0 && Y < (int)v.H;
    if (xin && yin) return view_set(v, X, Y);
    if (!v.gmask || (!xin && !yin)) return false;
    if (!xin) {
        const int x0 = X >= (int)v.W ? X - (int)v.W : X + (int)v.W;
        const bool ok = X >= (int)v.W ? x0 < v.xl : x0 > v.xr;
        return ok && x0 >= 0 && x0 < (int)v.W && px_set(v.gmask, v.tiles_x, x0, Y);
    }
    const int y0 = Y >= (int)v.H ? Y - (int)v.H : Y + (int)v.H;
    const bool ok = Y >= (int)v.H ? y0 < v.yb : y0 > v.yt;
    return ok && y0 >= 0 && y0 < (int)v.H && px_set(v.gmask, v.tiles_x, X, y0);
}

// Ghost flags of pixels (xs + i, Y), i = 0..7, as bits (0 outside [0, W)).
__device__ __forceinline__ uint32_t ghost_row(const KnnView& v, int Y, int xs) {
    const int c0 = xs >> 3, off = xs - c0 * 8, sh = (Y & 7) * 8;
    const uint32_t row = (uint32_t)(Y >> 3) * v.tiles_x;
    uint32_t b0 = 0, b1 = 0;
    if (c0 >= 0 && c0 < (int)v.tiles_x) b0 = (uint32_t)(gld(v.gmask + row + (uint32_t)c0) >> sh) & 0xffu;
    if (c0 + 1 >= 0 && c0 + 1 < (int)v.tiles_x) b1 = (uint32_t)(gld(v.gmask + row + (uint32_t)(c0 + 1)) >> sh) & 0xffu;
    return ((b0 | (b1 << 8)) >> off) & 0xffu;
}

// Bits i of [lo, hi) within 0..7.
__device__ __forceinline__ uint32_t bit_range(int lo, int hi) {
    lo = max(lo, 0);
    hi = min(hi, 8);
    return hi > lo ? ((1u << hi) - 1u) & ~((1u << lo) - 1u) : 0u;
}

// The points of 8x8 tile (gx, gy) as a tile word; in tiling mode the tile
// may lie partly or wholly outside the image and holds ghost copies there.
__device__ __forceinline__ uint64_t tile_points(const KnnView& v, int gx, int gy, bool& in_grid) {
    in_grid = gx >= 0 && gy >= 0 && gx < (int)v.tiles_x && gy < (int)v.tiles_y;
    uint64_t w = in_grid ? mask_word(v, (uint32_t)gy * v.tiles_x + (uint32_t)gx) : 0ull;
    const int X0 = gx * 8, Y0 = gy * 8, W = (int)v.W, H = (int)v.H;
    if (!v.gmask || (X0 >= 0 && Y0 >= 0 && X0 + 8 <= W && Y0 + 8 <= H)) return w;
    for (int j = 0; j < 8; ++j) {
        const int Y = Y0 + j;
        uint32_t bits = 0;
        if (Y >= 0 && Y < H) {
            if (X0 + 8 > W) bits |= ghost_row(v, Y, X0 - W) & bit_range(W - X0, v.xl + W - X0);
            if (X0 < 0) bits |= ghost_row(v, Y, X0 + W) & bit_range(v.xr + 1 - X0 - W, -X0);
        } else if (Y >= H) {
            if (Y - H < v.yb) bits = ghost_row(v, Y - H, X0) & bit_range(-X0, W - X0);
        } else if (Y + H > v.yt) {
            bits = ghost_row(v, Y + H, X0) & bit_range(-X0, W - X0);
        }
        w |= (uint64_t)bits << (j * 8);
    }
    return w;
}

// The pixels of tile (gx, gy)'s new-pixel word nw that were resolved before
// v.n_at; rank loads are issued eight at a time.
__device__ __forceinline__ uint64_t filter_new(const KnnView& v, int gx, int gy, uint64_t nw) {
    uint64_t keep = 0;
    while (nw) {
        int b[8];
        uint32_t r[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            b[q] = -1;
            if (nw) {
                b[q] = __ffsll((long long)nw) - 1;
                nw &= nw - 1;
            }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
            r[q] = b[q] >= 0 ? gld(v.rank + (uint32_t)(gy * 8 + (b[q] >> 3)) * v.W + (uint32_t)(gx * 8 + (b[q] & 7)))
                             : 0xffffffffu;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (b[q] >= 0 && r[q] < v.n_at) keep |= 1ull << b[q];
    }
    return keep;
}

// Exhaustive scan of the resolved list (sparse phase).
__device__ __forceinline__ void knn_list(const KnnView& a, WaveLds* L, uint32_t n_res, int x, int y, uint32_t k, int lane, Sel& s) {
    // four chunks of the list per memory round trip
    for (uint32_t base = 0; base < n_res; base += 4 * kWave) {
        uint32_t f[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t idx = base + u * kWave + lane;
            f[u] = idx < n_res ? gld(a.rlist + idx) : 0u;
        }
#pragma unroll 1
        for (int u = 0; u < 4; ++u) {
            const uint32_t idx = base + u * kWave + lane;
            if (base + u * kWave >= n_res) break;
            const bool v = idx < n_res;
            const int px = (int)(f[u] % a.W), py = (int)(f[u] / a.W);
            uint64_t key = ~0ull;
            if (v) key = make_key(px - x, py - y);
            sel_push(L->keys, s, k, v, key, lane);
            if (a.gmask) {
                // the point's ghost copies (flush_resolved order: x, then y)
                const bool g = v && px_set(a.gmask, a.tiles_x, px, py);
                const bool gx = g && (px < a.xl || px > a.xr);
                const bool gy = g && (py < a.yb || py > a.yt);
                const int sx = px < a.xl ? (int)a.W : -(int)a.W, sy = py < a.yb ? (int)a.H : -(int)a.H;
                sel_push(L->keys, s, k, gx, gx ? make_key(px + sx - x, py - y) : ~0ull, lane);
                sel_push(L->keys, s, k, gy, gy ? make_key(px - x, py + sy - y) : ~0ull, lane);
            }
        }
    }
}

// Integer square root of small non-negative t.
__device__ __forceinline__ int isqrt_small(int t) {
    int r = (int)sqrtf((float)t);
    while (r * r > t) --r;
    while ((r + 1) * (r + 1) <= t) ++r;
    return r;
}

// Tile offset of spiral index t: ring 0 is the query's tile; ring r >= 1
// holds the 8r tiles at Chebyshev distance r (top, right, bottom, left).
__device__ __forceinline__ void spiral_tile(int t, int& ox, int& oy) {
    if (t == 0) {
        ox = oy = 0;
        return;
    }
    const int r = (isqrt_small(t) + 1) >> 1;
    const int o = t - (2 * r - 1) * (2 * r - 1);
    const int side = o / (2 * r), off = o - side * (2 * r);
    if (side == 0) {
        ox = -r + off;
        oy = -r;
    } else if (side == 1) {
        ox = r;
        oy = -r + off;
    } else if (side == 2) {
        ox = r - off;
        oy = r;
    } else {
        ox = -r;
        oy = r - off;
    }
}

// Tiles each lane reads per round trip of the ring search. One (64 tiles a
// round trip) keeps the search's registers few: under the engine's 80-VGPR
// bound it spills 28 VGPRs, against 48 at 4 and 35 at 2; the 256-job mix ran 1.03, 0.98-1.00
// and 0.96 s at 4, 2 and 1, the plain 1024^2 step the same within 1%.
constexpr int kRingU = 1;

// Ring search over the 8x8-tile resolved mask: 64 tiles per step in spiral
// order, one tile per lane, the set bits of each tile taken lane-parallel.
// Stops when k collected points lie closer than any pixel outside the
// completed rings (every point that close has been collected).
__device__ __forceinline__ void knn_rings(const KnnView& a, WaveLds* L, int x, int y, uint32_t k, int lane, Sel& s,
                                          uint32_t& iters) {
    const int tx = x >> 3, ty = y >> 3;
    // tiles that can hold points: the grid, plus the ghost margins in tiling mode
    int g_x0 = 0, g_y0 = 0, g_x1 = (int)a.tiles_x - 1, g_y1 = (int)a.tiles_y - 1;
    if (a.gmask) {
        g_x0 = (1 - a.xl) >> 3;
        g_y0 = (1 - a.yb) >> 3;
        g_x1 = max(g_x1, ((int)a.W + a.xl - 1) >> 3);
        g_y1 = max(g_y1, ((int)a.H + a.yb - 1) >> 3);
    }
    const int rmax = max(max(tx - g_x0, g_x1 - tx), max(ty - g_y0, g_y1 - ty));
    const int t_end = (2 * rmax + 1) * (2 * rmax + 1);
    iters = 0;
    for (int t0 = 0; t0 < t_end; t0 += kRingU * kWave) {
        ++iters;
        // kRingU * 64 tiles per memory round trip, then their set bits lane-parallel
        uint64_t word[kRingU], nword[kRingU];
        uint32_t tmin[kRingU];
        int gx[kRingU], gy[kRingU];
#pragma unroll
        for (int u = 0; u < kRingU; ++u) {
            const int t = t0 + u * kWave + lane;
            word[u] = nword[u] = 0;
            tmin[u] = 0xffffffffu;
            gx[u] = gy[u] = 0;
            if (t < t_end) {
                int ox, oy;
                spiral_tile(t, ox, oy);
                gx[u] = tx + ox;
                gy[u] = ty + oy;
                if (gx[u] >= g_x0 && gy[u] >= g_y0 && gx[u] <= g_x1 && gy[u] <= g_y1) {
                    bool in_grid;
                    word[u] = tile_points(a, gx[u], gy[u], in_grid);
                    if (a.mask_new && in_grid) {
                        const uint32_t ti = (uint32_t)gy[u] * a.tiles_x + (uint32_t)gx[u];
                        nword[u] = gld(a.mask_new + ti);
                        tmin[u] = gld(a.tile_min + ti);
                    }
                }
            }
        }
        if (a.mask_new) {
#pragma unroll
            for (int u = 0; u < kRingU; ++u)
                if (nword[u] && tmin[u] < a.n_at) word[u] |= filter_new(a, gx[u], gy[u], nword[u]);
        }
#pragma unroll 1
        for (int u = 0; u < kRingU; ++u) {
            uint64_t rest = word[u];
            while (__ballot(rest != 0)) {
                const bool have = rest != 0;
                uint64_t key = ~0ull;
                if (have) {
                    const int b = __ffsll((long long)rest) - 1;
                    rest &= rest - 1;
                    key = make_key(gx[u] * 8 + (b & 7) - x, gy[u] * 8 + (b >> 3) - y);
                }
                sel_push(L->keys, s, k, have, key, lane);
            }
        }
        // rings fully visited after these tiles
        const int rc = (isqrt_small(min(t0 + kRingU * kWave, t_end)) - 1) >> 1;
        if (rc >= rmax) break;
        if (s.count >= k && rc >= 0) {
            int64_t D = INT64_MAX, e;
            if (tx - rc > g_x0 && (e = x - (tx - rc) * 8 + 1) < D) D = e;
            if (tx + rc < g_x1 && (e = (tx + rc + 1) * 8 - x) < D) D = e;
            if (ty - rc > g_y0 && (e = y - (ty - rc) * 8 + 1) < D) D = e;
            if (ty + rc < g_y1 && (e = (ty + rc + 1) * 8 - y) < D) D = e;
            const uint64_t lim = (uint64_t)(D * D) << 32;
            uint32_t c = 0;
            for (int q = lane; q < (int)s.count; q += kWave) c += L->keys[q] < lim;
            for (int off = kWave / 2; off > 0; off >>= 1) c += (uint32_t)__shfl_xor((int)c, off);
            if (c >= k) break;
        }
    }
}

// Ring-ordered offset table scan: the first k resolved pixels in table
// order are the k nearest (the table lists every offset within its radius
// in canonical order). KU x 64 offsets per memory round trip, taken in table
// order: 4 (256) where latency binds; fewer where the VALU binds, since a
// dense mask holds k neighbours within the first 64 or so. Returns the
// number found (<= k).
template <int KU>
__device__ __forceinline__ uint32_t knn_table(const KnnView& a, WaveLds* L, int x, int y, uint32_t k, int lane) {
    uint32_t found = 0;
    for (uint32_t base = 0; base < a.offtab_n && found < k; base += KU * kWave) {
        int32_t o[KU];
        bool hit[KU];
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            const uint32_t t = base + u * kWave + lane;
            o[u] = 0;
            hit[u] = false;
            if (t < a.offtab_n) {
                o[u] = (a.offc && t < kOffCache) ? a.offc[t] : gld(a.offtab + t);
                const int X = x + packed_dx(o[u]), Y = y + packed_dy(o[u]);
                const bool in = X >= 0 && Y >= 0 && X < (int)a.W && Y < (int)a.H;
                hit[u] = in ? view_set(a, X, Y) : point_at(a, X, Y);
                if (in && !hit[u] && a.mask_new && px_set(a.mask_new, a.tiles_x, X, Y))
                    hit[u] = gld_at(a.rank, (__umul24((uint32_t)Y, a.W) + (uint32_t)X) * 4u) < a.n_at;
            }
        }
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            const uint64_t m = __ballot(hit[u]);
            const uint32_t pre = bits_below(m);
            if (hit[u] && found + pre < k) L->nd[found + pre] = o[u];
            found += (uint32_t)__popcll(m);
        }
    }
    return found < k ? found : k;
}

// k nearest of (x, y) among the view's resolved pixels (n_res of them), in
// canonical order, into L->nd; returns the count (<= k). KU: the offset
// table's offsets per lane per round trip.
template <int KU = 4>
__device__ __forceinline__ uint32_t knn_search(const KnnView& v, WaveLds* L, int x, int y, uint32_t k, uint32_t n_res,
                                               int lane, uint32_t& iters, uint32_t& buffered) {
    uint32_t found = 0;
    bool use_sel = n_res <= kBruteMax;
    // the offset table pays off when about k / density <= 1024 offsets (four
    // round trips) reach k resolved pixels; sparser masks use the ring search
    const bool sparse = (uint64_t)n_res * 1024ull < (uint64_t)k * v.W * v.H;
    if (!use_sel && !sparse) {
        found = knn_table<KU>(v, L, x, y, k, lane);
        // fewer than k within the table: search on, unless every point was
        // found (without ghosts there are n_res of them)
        use_sel = found < k && (v.gmask || found < n_res);
    } else if (!use_sel) {
        use_sel = true;
    }
    iters = buffered = 0;
    if (use_sel) {
        Sel s{0u, ~0ull};
        if (n_res <= kBruteMax) knn_list(v, L, n_res, x, y, k, lane, s);
        else knn_rings(v, L, x, y, k, lane, s, iters);
        buffered = s.count;
        wave_sync();
        s = sel_compact(L->keys, s, k, lane);
        found = s.count;
        if (lane < (int)found) L->nd[lane] = key_to_packed(L->keys[lane]);
    }
    wave_sync();
    return found;
}

__device__ __forceinline__ bool coord_ok(const MapDesc& md, int x, int y) {
    if (x < 0 || y < 0 || (uint32_t)x >= md.w || (uint32_t)y >= md.h) return false;
    return md.smask == nullptr || gld(md.smask + (uint32_t)y * md.w + (uint32_t)x) != 0;
}

// Byte c of v times 2^SH: one SDWA shift. The compiler turns byte 0 into a
// shift and a mask, so that one is written out (a pure VALU instruction).
template <int c, int SH>
__device__ __forceinline__ uint32_t byte_scaled(uint32_t v) {
    if constexpr (c == 0) {
        uint32_t r;
        asm("v_lshlrev_b32_sdwa %0, %2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_0"
            : "=v"(r)
            : "v"(v), "I"(SH));
        return r;
    } else {
        return ((v >> (8 * c)) & 255u) << SH;
    }
}

// Cost tables are interleaved C ways in LDS: entry d of copy s is word
// d*C + s. Channel bytes scaled by the entry stride (4C bytes) give an
// entry's byte offset from their absolute difference, one v_sad_u32 with
// the copy's byte offset slot4 as the addend.
template <int C>
constexpr int lut_shift() {
    static_assert(C == 8 || C == 4, "table copies");
    return C == 8 ? 5 : 4;
}
__device__ __forceinline__ float lut_sad(const float* lut, uint32_t a, uint32_t b, uint32_t slot4) {
    const uint32_t off = (a > b ? a - b : b - a) + slot4;
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(lut) + off);
}

// The four channel bytes of v, scaled for lut_sad.
template <int C>
__device__ __forceinline__ uint4 bytes_scaled(uint32_t v) {
    constexpr int SH = lut_shift<C>();
    return make_uint4(byte_scaled<0, SH>(v), byte_scaled<1, SH>(v), byte_scaled<2, SH>(v), byte_scaled<3, SH>(v));
}

// s plus the four channel terms |q_c - p_c| of table `lut`, added in R, G, B,
// A order; q is given scaled (bytes_scaled).
template <int C>
__device__ __forceinline__ float lut4_add(float s, const float* lut, const uint4& q, uint32_t p, uint32_t slot4) {
    constexpr int SH = lut_shift<C>();
    s = s + lut_sad(lut, q.x, byte_scaled<0, SH>(p), slot4);
    s = s + lut_sad(lut, q.y, byte_scaled<1, SH>(p), slot4);
    s = s + lut_sad(lut, q.z, byte_scaled<2, SH>(p), slot4);
    s = s + lut_sad(lut, q.w, byte_scaled<3, SH>(p), slot4);
    return s;
}

// The four channel terms alone (the first one starts the sum).
template <int C>
__device__ __forceinline__ float lut4(const float* lut, const uint4& q, uint32_t p, uint32_t slot4) {
    constexpr int SH = lut_shift<C>();
    float s = lut_sad(lut, q.x, byte_scaled<0, SH>(p), slot4);
    s = s + lut_sad(lut, q.y, byte_scaled<1, SH>(p), slot4);
    s = s + lut_sad(lut, q.z, byte_scaled<2, SH>(p), slot4);
    s = s + lut_sad(lut, q.w, byte_scaled<3, SH>(p), slot4);
    return s;
}

// Candidate ci of the reference's list: valid coherence candidates in
// neighbour order, then the random ones (find_candidates, ms.rs:496-599).
__device__ __forceinline__ void cand_get(const BatchArgs& a, const WaveLds* L, uint32_t rel, uint32_t item,
                                         uint32_t n_coh, uint32_t ci, uint32_t& xy, uint32_t& ms) {
    if (a.ext_cands) {
        uint4 e = a.ext_cands[(size_t)rel * a.ext_cmax + ci];
        xy = e.x;
        ms = e.y;
    } else if (ci < n_coh) {
        uint4 e = L->coh[ci];
        xy = e.x;
        ms = e.y;
    } else {
        uint2 r = gld(a.rcands + (size_t)(item - a.rc_base) * a.m + (ci - n_coh));
        xy = r.x;
        ms = r.y | 0x80000000u;
    }
}

// Gather of a candidate's pattern pixel for neighbour j (mirrored for
// random candidates, ms.rs:534/588): colour and, when guided, guide pixel;
// out-of-bounds pixels are [0,0,0,255] (get_color_of_neighbor, ms.rs:1122-1149).
// Pattern pixel of a candidate at xy for a neighbour at (dx, dy), the
// offset already negated for a mirrored pattern. The descriptor's fields are
// read together, ahead of the bounds test, so that their loads are not
// chained behind one another.
template <bool GUIDED>
__device__ __forceinline__ void gather_xy(const MapDesc& md, uint32_t xy, int dx, int dy, uint32_t& p, uint32_t& gp) {
    const uint32_t* px = md.px;
    const uint32_t w = md.w, h = md.h;
    const int X2 = (int)(xy & 0xffffu) + dx, Y2 = (int)(xy >> 16) + dy;
    // in bounds, Y2 and w are below 2^16 and a level holds at most 2^30
    // pixels (checked at job creation), so the byte offset fits 32 bits
    const bool in = (uint32_t)X2 < w && (uint32_t)Y2 < h;
    p = in ? gld_at(px, (__umul24((uint32_t)Y2, w) + (uint32_t)X2) * 4u) : kOutsidePx;
    gp = kOutsidePx;
    if (GUIDED) {
        const uint32_t* gpx = md.gpx;
        const uint32_t gw = md.gw, gh = md.gh;
        const bool gin = (uint32_t)X2 < gw && (uint32_t)Y2 < gh;
        gp = gin ? gld_at(gpx, (__umul24((uint32_t)Y2, gw) + (uint32_t)X2) * 4u) : kOutsidePx;
    }
}
// Candidate (xy, map word ms): its map, and its mirror bit applied per lane.
template <bool GUIDED>
__device__ __forceinline__ void gather_at(const MapDesc* maps, uint32_t xy, uint32_t ms, int dx, int dy,
                                          uint32_t& p, uint32_t& gp) {
    const bool mir = (ms >> 31) != 0;
    gather_xy<GUIDED>(maps[ms & 0x7fffffffu], xy, mir ? -dx : dx, mir ? -dy : dy, p, gp);
}

// Target side of neighbour j, the same for every candidate: its pattern and
// guide pixels scaled for lut_sad, its offset and its weight.
struct NbrRow {
    uint4 q, g;
    int dx, dy;
    float w;
};
template <bool GUIDED>
__device__ __forceinline__ NbrRow nbr_row(const WaveLds* L, uint32_t j) {
    constexpr int C = lut_copies<GUIDED>();
    NbrRow r;
    r.q = bytes_scaled<C>(L->pat[j]);
    r.g = GUIDED ? bytes_scaled<C>(L->gpat[j]) : make_uint4(0, 0, 0, 0);
    const int32_t o = L->nd[j];
    r.dx = packed_dx(o);
    r.dy = packed_dy(o);
    r.w = L->w[j];
    return r;
}

// Cost term of a neighbour: (four colour LUT terms, then four guide LUT
// terms, summed in order) times w_j, as one rounded product.
template <bool GUIDED>
__device__ __forceinline__ float term_at(const NbrRow& r, const float* lut, const float* glut, uint32_t slot4,
                                         uint32_t p, uint32_t gp) {
    constexpr int C = lut_copies<GUIDED>();
    float s = lut4<C>(lut, r.q, p, slot4);
    if (GUIDED) s = lut4_add<C>(s, glut, r.g, gp, slot4);
    return __fmul_rn(s, r.w);
}

struct Best {
    float B;      // lowest complete cost so far (FLT_MAX: none)
    uint32_t Bi;  // its candidate index (kNone: none)
    uint32_t xy, ms;  // its coordinate and map word (bit 31: mirrored pattern)
};

// Evaluates the n_act candidates in L->act_* with exact pruning. Terms of a
// chunk of neighbours are computed with one lane per (candidate, neighbour)
// pair; each candidate then adds its terms in neighbour order (the
// reference's sequential f32 sum). A candidate whose partial sum already
// exceeds the best complete cost (or equals it with a later index) can
// never win and is dropped: partial sums of non-negative terms only grow.
// The pruning bound: the lexicographically smaller of the member's own best
// and the team's (read by one lane so the wave sees one value).
__device__ __forceinline__ void prune_bound(const unsigned long long* tb, const Best& best, float& PB, uint32_t& PBi) {
    PB = best.B;
    PBi = best.Bi;
    if (!tb) return;
    const unsigned long long v = __hip_atomic_load(tb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const uint32_t vi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t vb = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    const float B = __uint_as_float(vb);
    if (vi != kNone && (PBi == kNone || B < PB || (B == PB && vi < PBi))) {
        PB = B;
        PBi = vi;
    }
}

// FIRST_SH: while a pruning bound exists, the first chunk holds 2^FIRST_SH
// neighbours. 3 (8 neighbours) starts pruning early, which pays where the
// VALU binds (the multi-job engine); 6 (a full chunk) saves a gather round
// trip, which pays where latency binds (the single-job kernels). A constant
// here also lets the compiler keep the loop in fewer registers.
template <bool GUIDED, int FIRST_SH>
__device__ __forceinline__ void eval_group(const MapDesc* maps, uint32_t n_maps, WaveLds* L, const float* lut, const float* glut,
                                           int slot, int lane, uint32_t kf, uint32_t n_act, Best& best,
                                           unsigned long long* tb) {
    const uint32_t slot4 = (uint32_t)slot * 4u;
    MapDesc m0{};
    if (n_maps == 1) m0 = ld_uniform(maps);
    if (lane < (int)n_act) L->act_sum[lane] = 0.0f;
    wave_sync();
    // the group's mirror state: 0 none, 1 all, 2 mixed. A group holds
    // coherence candidates only (never mirrored) or random ones only (always
    // mirrored); host-given candidate lists may mix them.
    const uint64_t g_all = __ballot(lane < (int)n_act);
    const uint64_t g_mm = __ballot(lane < (int)n_act && (L->act_ms[lane] >> 31) != 0);
    const int g_mir = g_mm == 0 ? 0 : (g_mm == g_all ? 1 : 2);
    float PB;
    uint32_t PBi;
    prune_bound(tb, best, PB, PBi);
    uint32_t j0 = 0;
    while (n_act > 0) {
        // chunk of 2^sh neighbours with n_act << sh <= kTermCap; while a bound
        // exists the first chunk holds 2^FIRST_SH
        uint32_t sh = (j0 == 0 && PBi != kNone) ? (uint32_t)FIRST_SH : 6u;
        while (sh > 0 && (n_act << sh) > (uint32_t)kTermCap) --sh;
        const uint32_t jc = 1u << sh;
        const uint32_t jn = min(jc, kf - j0);
        const uint32_t total = n_act << sh;
        // jc divides the wave, so a lane's neighbour is the same in every
        // slice u of the chunk: its row is read once
        const uint32_t jj = (uint32_t)lane & (jc - 1);
        const bool jin = jj < jn;
        const NbrRow row = nbr_row<GUIDED>(L, j0 + (jin ? jj : 0u));
        // all gathers of the chunk first (at most kTermCap / 64 per lane), then
        // the LUT sums. With one example (its descriptor held in scalars) and
        // one mirror state for the group, the offset's sign is applied once.
        uint32_t px[kTermCap / kWave], gpx[kTermCap / kWave];
        if (n_maps == 1 && g_mir != 2) {
            const int ddx = g_mir ? -row.dx : row.dx, ddy = g_mir ? -row.dy : row.dy;
#pragma unroll
            for (int u = 0; u < kTermCap / kWave; ++u) {
                const uint32_t l = (uint32_t)(u * kWave + lane), ai = l >> sh;
                px[u] = gpx[u] = 0;
                if (l < total && jin) gather_xy<GUIDED>(m0, L->act_xy[ai], ddx, ddy, px[u], gpx[u]);
            }
        } else {
#pragma unroll
            for (int u = 0; u < kTermCap / kWave; ++u) {
                const uint32_t l = (uint32_t)(u * kWave + lane), ai = l >> sh;
                px[u] = gpx[u] = 0;
                if (l < total && jin) gather_at<GUIDED>(maps, L->act_xy[ai], L->act_ms[ai], row.dx, row.dy, px[u], gpx[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < kTermCap / kWave; ++u) {
            const uint32_t l = (uint32_t)(u * kWave + lane);
            if (l < total && jin) L->terms[l] = term_at<GUIDED>(row, lut, glut, slot4, px[u], gpx[u]);
        }
        wave_sync();
        const bool act = lane < (int)n_act;
        float s = 0.0f;
        uint32_t myidx = kNone, xy = 0, ms = 0;
        if (act) {
            s = L->act_sum[lane];
            // the terms are loaded 16 at a time ahead of the in-order adds
            const float* tp = &L->terms[lane << sh];
            uint32_t jj = 0;
            if (sh >= 2) {
                for (; jj + 16 <= jn; jj += 16) {
                    const float4 q0 = *reinterpret_cast<const float4*>(tp + jj);
                    const float4 q1 = *reinterpret_cast<const float4*>(tp + jj + 4);
                    const float4 q2 = *reinterpret_cast<const float4*>(tp + jj + 8);
                    const float4 q3 = *reinterpret_cast<const float4*>(tp + jj + 12);
                    s = __fadd_rn(s, q0.x); s = __fadd_rn(s, q0.y); s = __fadd_rn(s, q0.z); s = __fadd_rn(s, q0.w);
                    s = __fadd_rn(s, q1.x); s = __fadd_rn(s, q1.y); s = __fadd_rn(s, q1.z); s = __fadd_rn(s, q1.w);
                    s = __fadd_rn(s, q2.x); s = __fadd_rn(s, q2.y); s = __fadd_rn(s, q2.z); s = __fadd_rn(s, q2.w);
                    s = __fadd_rn(s, q3.x); s = __fadd_rn(s, q3.y); s = __fadd_rn(s, q3.z); s = __fadd_rn(s, q3.w);
                }
                for (; jj + 4 <= jn; jj += 4) {
                    const float4 q = *reinterpret_cast<const float4*>(tp + jj);
                    s = __fadd_rn(s, q.x); s = __fadd_rn(s, q.y); s = __fadd_rn(s, q.z); s = __fadd_rn(s, q.w);
                }
            }
            for (; jj < jn; ++jj) s = __fadd_rn(s, tp[jj]);
            myidx = L->act_idx[lane];
            xy = L->act_xy[lane];
            ms = L->act_ms[lane];
        }
        j0 += jn;
        if (j0 >= kf) {
            float key = act ? s : INFINITY;
            uint32_t ki = myidx, kxy = xy, kms = ms;
            for (int off = kWave / 2; off > 0; off >>= 1) {
                float ok = __shfl_xor(key, off);
                uint32_t oi = (uint32_t)__shfl_xor((int)ki, off);
                uint32_t oxy = (uint32_t)__shfl_xor((int)kxy, off);
                uint32_t oms = (uint32_t)__shfl_xor((int)kms, off);
                if (ok < key || (ok == key && oi < ki)) {
                    key = ok;
                    ki = oi;
                    kxy = oxy;
                    kms = oms;
                }
            }
            if (key < best.B || (key == best.B && best.Bi != kNone && ki < best.Bi)) {
                best.B = key;
                best.Bi = ki;
                best.xy = kxy;
                best.ms = kms;
            }
            // publish: another member's complete cost bounds this one's candidates
            if (tb && lane == 0 && best.Bi != kNone)
                __hip_atomic_fetch_min(tb, ((unsigned long long)__float_as_uint(best.B) << 32) | best.Bi,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            break;
        }
        prune_bound(tb, best, PB, PBi);
        const bool keep = act && !(s > PB || (s == PB && (PBi == kNone || myidx > PBi)));
        const uint64_t km = __ballot(keep);
        const uint32_t pos = bits_below(km);
        wave_sync();
        if (keep) {
            L->act_xy[pos] = xy;
            L->act_ms[pos] = ms;
            L->act_idx[pos] = myidx;
            L->act_sum[pos] = s;
        }
        n_act = (uint32_t)__popcll(km);
        wave_sync();
    }
    wave_sync();
}

__device__ __forceinline__ void phase_clock(const BatchArgs& a, uint32_t item, int phase, int lane) {
    if (a.dbg_clk && lane == 0) a.dbg_clk[(size_t)item * 16 + phase] = wall_clock64();
}
__device__ __forceinline__ void phase_value(const BatchArgs& a, uint32_t item, int slot, uint64_t v, int lane) {
    if (a.dbg_clk && lane == 0) a.dbg_clk[(size_t)item * 16 + slot] = v;
}

// ---- one work item, resolved by a team of T wavefronts (T = 1, 2, 4, 8, 16) ----
//
// The team leader (member 0) finds the k nearest resolved neighbours; every
// member then gathers the neighbour state, weights and coherence candidates
// itself (identical values), evaluates its own slice of the candidate list
// with exact pruning, and leaves its best (cost, index) in LDS; the leader
// takes the lexicographic minimum, which is the reference's winner (the
// first candidate of least cost), and writes the result. Members of one
// team share a workgroup; the two workgroup barriers are taken by every wave
// of the workgroup whatever its item does, so T > 1 needs all waves of the
// workgroup to call this together. With T = 1 there is no barrier.

// Leader, step 1: kNN into L->nd; returns the count. Test hooks: host lists,
// dumps, knn_only.
template <int KU>
__device__ __forceinline__ uint32_t item_knn(const BatchArgs& a, uint32_t item, uint32_t rel, int x, int y,
                                             uint32_t n_res, WaveLds* L, int lane) {
    const uint32_t k = a.k;
    uint32_t kf;
    if (a.ext_nbrs) {
        kf = a.ext_nbr_count[rel];
        if (lane < (int)kf) L->nd[lane] = a.ext_nbrs[(size_t)rel * k + lane];
    } else if (a.pre_nbrs && item >= a.pre_base && item < a.pre_end) {
        kf = gld(a.pre_cnt + (item - a.pre_base));
        if (lane < (int)kf) L->nd[lane] = gld(a.pre_nbrs + (size_t)(item - a.pre_base) * k + lane);
    } else {
        const KnnView v{a.W,     a.H,     a.tiles_x, a.tiles_y, a.mask, a.rlist, a.offtab, a.offtab_n, nullptr,
                        nullptr, nullptr, 0u,        a.gmask,   a.xl,   a.xr,    a.yb,     a.yt,       a.pend_r,
                        offtab_lds()};
        uint32_t iters, buffered;
        kf = knn_search<KU>(v, L, x, y, k, n_res, lane, iters, buffered);
        phase_clock(a, item, 7, lane);
        phase_value(a, item, 8, iters, lane);
        phase_value(a, item, 9, buffered, lane);
    }
    wave_sync();
    if (a.dbg_nbrs) {
        if (lane < (int)kf) a.dbg_nbrs[(size_t)rel * k + lane] = L->nd[lane];
        if (lane == 0) a.dbg_nbr_count[rel] = kf;
    }
    return kf;
}

// Leader: no resolved neighbour, resolve_at_random with the stage seed
// (ms.rs:1003-1009). Returns the result (lane 0), also written to
// result[rel] unless a.result is null.
__device__ __forceinline__ uint4 item_at_random(const BatchArgs& a, const MapDesc* maps, uint32_t item, uint32_t rel,
                                                uint32_t flat, int lane) {
    if (lane != 0) return make_uint4(0, 0, 0, 0);
    Pcg32 g0 = Pcg32::seed_from_u64(a.stage_seed);
    uint32_t rm = (uint32_t)range_u64(g0, 0, a.n_maps);
    Pcg32 g1 = Pcg32::seed_from_u64(a.stage_seed);
    uint32_t rx = range_u32(g1, 0, maps[rm].w);
    Pcg32 g2 = Pcg32::seed_from_u64(a.stage_seed);
    uint32_t ry = range_u32(g2, 0, maps[rm].h);
    const uint4 r = make_uint4(rx | (ry << 16), rm | (rm << 16) | kFallbackBit, flat, __float_as_uint(0.0f));
    if (a.result) gst(a.result + rel, r);
    if (a.item_score) gst(a.item_score + item, -1.0f);
    if (a.dbg_best) a.dbg_best[rel] = 0xffffffffu;
    return r;
}

// Every member: neighbour state (own pattern, distances, weights) into L and
// the coherence candidates; returns n_coh. Lanes < n_k0 hold candidate
// lane's (xy, map) in xy/ms; `uniq` marks the first of equal ones.
// The neighbour weights (ms.rs:405-425, 1198-1203), from the neighbour
// offsets alone: distances, their average over the 4k duplicated entries
// summed in order, exp(-d / avg).
__device__ __forceinline__ void item_geometry(const BatchArgs& a, int x, int y, uint32_t kf, WaveLds* L, int lane) {
    if (lane < (int)kf) {
        const int32_t o = L->nd[lane];
        const int nx = x + packed_dx(o), ny = y + packed_dy(o);
        const double dimx = (double)a.W, dimy = (double)a.H;
        const double x1 = (double)nx / dimx, y1 = (double)ny / dimy;
        const double x2 = (double)x / dimx, y2 = (double)y / dimy;
        L->dist[lane] = fma(x1 - x2, x1 - x2, (y1 - y2) * (y1 - y2));
    }
    wave_sync();
    if (lane == 0) {
        // the reference's sum over the 4k duplicated distances, in order; the
        // distances are read eight at a time ahead of the adds
        double sum = 0.0;
        uint32_t j = 0;
        for (; j + 8 <= kf; j += 8) {
            const double2 d0 = *reinterpret_cast<const double2*>(&L->dist[j]);
            const double2 d1 = *reinterpret_cast<const double2*>(&L->dist[j + 2]);
            const double2 d2 = *reinterpret_cast<const double2*>(&L->dist[j + 4]);
            const double2 d3 = *reinterpret_cast<const double2*>(&L->dist[j + 6]);
            const double dd[8] = {d0.x, d0.y, d1.x, d1.y, d2.x, d2.y, d3.x, d3.y};
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                sum = sum + dd[q];
                sum = sum + dd[q];
                sum = sum + dd[q];
                sum = sum + dd[q];
            }
        }
        for (; j < kf; ++j) {
            const double d = L->dist[j];
            sum = sum + d;
            sum = sum + d;
            sum = sum + d;
            sum = sum + d;
        }
        L->avg = sum / (double)(4u * kf);
    }
    wave_sync();
    if (lane < (int)kf) L->w[lane] = (float)exp_dd(-1.0 * (L->dist[lane] / L->avg));
    wave_sync();
}

// old_state (redo runs): neighbours without a bit in `fresh` are read there.
template <bool GUIDED>
__device__ __forceinline__ uint32_t item_neighbours(const BatchArgs& a, const MapDesc* maps, uint32_t item,
                                                    uint32_t rel, int x, int y, uint32_t kf, uint32_t n_k0,
                                                    WaveLds* L, int lane, bool geo_done,
                                                    uint32_t& xy, uint32_t& ms, uint64_t& uniq,
                                                    const uint4* old_state = nullptr, uint64_t fresh = 0) {
    bool coh_valid = false;
    uint4 coh = make_uint4(0, 0, 0, 0);
    if (lane < (int)kf) {
        int32_t o = L->nd[lane];
        int nx = x + packed_dx(o), ny = y + packed_dy(o);
        if (a.gmask) {
            // a ghost neighbour reads the pixel it copies (ms.rs:510-512, 1132-1136)
            nx = nx < 0 ? nx + (int)a.W : (nx >= (int)a.W ? nx - (int)a.W : nx);
            ny = ny < 0 ? ny + (int)a.H : (ny >= (int)a.H ? ny - (int)a.H : ny);
        }
        uint32_t nflat = (uint32_t)ny * a.W + (uint32_t)nx;
        uint4 st;
        if (a.state2) {
            // deferred commit: the previous batch's records are still in state2
            const uint32_t sp = gld(a.stamp + nflat);
            const uint4 s0 = gld(a.state + nflat), s1 = gld(a.state2 + nflat);
            st = sp == a.serial - 1u ? s1 : s0;
        } else {
            const uint4* src = (old_state && !((fresh >> lane) & 1ull)) ? old_state : a.state;
            st = gld(src + nflat);
        }
        L->pat[lane] = st.x;
        if (GUIDED) L->gpat[lane] = gld(a.tg + nflat);
        if (!a.ext_cands) {
            int cx = (int)(st.y & 0xffffu) - packed_dx(o);
            int cy = (int)(st.y >> 16) - packed_dy(o);
            uint32_t pmap = st.z >> 16;
            coh_valid = pmap < a.n_maps && coord_ok(maps[pmap], cx, cy);
            coh = make_uint4((uint32_t)cx | ((uint32_t)cy << 16), pmap, st.w, 0u);
        }
    }
    uint64_t cm = __ballot(coh_valid);
    if (coh_valid) L->coh[bits_below(cm)] = coh;
    const uint32_t n_coh = (uint32_t)__popcll(cm);
    wave_sync();
    if (!geo_done) item_geometry(a, x, y, kf, L, lane);
    // keep the first of equal (coordinate, map) coherence candidates: equal
    // patterns give equal costs and the first index wins ties (a host list
    // leads with its n_k0 coherence entries)
    const uint32_t nk = a.ext_cands ? n_k0 : n_coh;
    xy = 0;
    ms = 0;
    const bool v = lane < (int)nk;
    if (v) cand_get(a, L, rel, item, n_coh, (uint32_t)lane, xy, ms);
    uint64_t pending = __ballot(v);
    uniq = 0;
    while (pending) {
        int l = __ffsll((long long)pending) - 1;
        uniq |= 1ull << l;
        uint32_t kxy = (uint32_t)__shfl((int)xy, l), kms = (uint32_t)__shfl((int)ms, l);
        pending &= ~__ballot(v && xy == kxy && ms == kms);
    }
    return n_coh;
}

// Leader: every candidate without pruning, find_best_match's sequential
// rule, for NaN cost tables and the full-cost test hook.
template <bool GUIDED>
__device__ __forceinline__ void item_exact(const BatchArgs& a, const MapDesc* maps, uint32_t item, uint32_t rel,
                                           uint32_t kf, uint32_t n_coh, uint32_t n_c, WaveLds* L, const float* lut,
                                           const float* glut, int slot, int lane, float& best, uint32_t& bidx) {
    const uint32_t slot4 = (uint32_t)slot * 4u;
    for (uint32_t base = 0; base < n_c; base += kWave) {
        const uint32_t c = base + lane;
        const bool v = c < n_c;
        float F = 0.0f, X = -INFINITY;
        if (v) {
            uint32_t cxy, cmap;
            int sgn;
            if (a.ext_cands) {
                uint4 e = a.ext_cands[(size_t)rel * a.ext_cmax + c];
                cxy = e.x;
                cmap = e.y & 0x7fffffffu;
                sgn = (e.y >> 31) ? -1 : 1;
            } else if (c < n_coh) {
                uint4 e = L->coh[c];
                cxy = e.x;
                cmap = e.y;
                sgn = 1;
            } else {
                uint2 r = gld(a.rcands + (size_t)(item - a.rc_base) * a.m + (c - n_coh));
                cxy = r.x;
                cmap = r.y;
                sgn = -1;
            }
            const MapDesc md = maps[cmap];
            const int cx = (int)(cxy & 0xffffu), cy = (int)(cxy >> 16);
            float score = 0.0f, lastgood = -INFINITY;
            for (uint32_t j = 0; j < kf; ++j) {
                const int32_t o = L->nd[j];
                const int X2 = cx + sgn * packed_dx(o), Y2 = cy + sgn * packed_dy(o);
                const bool in = (uint32_t)X2 < md.w && (uint32_t)Y2 < md.h;
                const uint32_t idx = (uint32_t)Y2 * md.w + (uint32_t)X2;
                const uint32_t p = in ? gld(md.px + idx) : kOutsidePx;
                constexpr int C = lut_copies<GUIDED>();
                float s = lut4<C>(lut, bytes_scaled<C>(L->pat[j]), p, slot4);
                if (GUIDED) {
                    const bool gin = (uint32_t)X2 < md.gw && (uint32_t)Y2 < md.gh;
                    const uint32_t gp = gin ? md.gpx[(uint32_t)Y2 * md.gw + (uint32_t)X2] : kOutsidePx;
                    s = lut4_add<C>(s, glut, bytes_scaled<C>(L->gpat[j]), gp, slot4);
                }
                score = __fadd_rn(score, __fmul_rn(s, L->w[j]));
                if (score == score) lastgood = score;
            }
            F = score;
            X = (F == F) ? F : lastgood;
            if (a.ext_full_cost) a.ext_full_cost[(size_t)rel * a.ext_cmax + c] = F;
        }
        // find_best_match's sequential rule: candidate c replaces the best
        // iff no partial sum reached it, i.e. !(X_c >= best) (partial sums of
        // non-negative terms are monotone; X_c is the last non-NaN partial).
        const bool anynan = __ballot(v && !(F == F)) != 0ull;
        if (!anynan && best == best) {
            float key = v ? F : INFINITY;
            uint32_t ki = c;
            for (int off = kWave / 2; off > 0; off >>= 1) {
                float ok = __shfl_xor(key, off);
                uint32_t oi = (uint32_t)__shfl_xor((int)ki, off);
                if (ok < key || (ok == key && oi < ki)) {
                    key = ok;
                    ki = oi;
                }
            }
            if (key < best) {
                best = key;
                bidx = ki;
            }
        } else {
            for (int l = 0; l < kWave; ++l) {
                if (base + (uint32_t)l >= n_c) break;
                float Xl = __shfl(X, l), Fl = __shfl(F, l);
                if (!(Xl >= best)) {
                    best = Fl;
                    bidx = base + (uint32_t)l;
                }
            }
        }
    }
}

// Leader: the winning candidate. Returns the result (lane 0), also written
// to result[rel] unless a.result is null. xm: the winner's coordinate and
// map word when the evaluation carried them (else they are read again).
__device__ __forceinline__ uint4 item_winner(const BatchArgs& a, const MapDesc* maps, uint32_t item, uint32_t rel,
                                             uint32_t n_coh, const WaveLds* L, float best, uint32_t bidx, int lane,
                                             const uint2* xm = nullptr) {
    if (lane != 0) return make_uint4(0, 0, 0, 0);
    uint32_t cxy, cmap, patch;
    if (a.ext_cands) {
        uint4 e = a.ext_cands[(size_t)rel * a.ext_cmax + bidx];
        cxy = e.x;
        cmap = e.y & 0x7fffffffu;
        patch = e.z;
    } else if (bidx < n_coh) {
        uint4 e = L->coh[bidx];
        cxy = e.x;
        cmap = e.y;
        patch = e.z;
    } else {
        if (xm) {
            cxy = xm->x;
            cmap = xm->y & 0x7fffffffu;
        } else {
            const uint2 rc = gld(a.rcands + (size_t)(item - a.rc_base) * a.m + (bidx - n_coh));
            cxy = rc.x;
            cmap = rc.y;
        }
        patch = (cxy >> 16) * maps[cmap].w + (cxy & 0xffffu);
    }
    const uint4 r = make_uint4(cxy, cmap | (cmap << 16), patch, __float_as_uint(best));
    if (a.result) gst(a.result + rel, r);
    if (a.item_score) gst(a.item_score + item, best);
    if (a.dbg_best) a.dbg_best[rel] = bidx;
    return r;
}

// Resolves work item `item` of the batch starting at item0 against the state
// frozen at the batch start (n_res resolved pixels); writes result[item -
// item0]. Lt: the team's WaveLds (member q uses Lt + q); `valid` is false
// for members of a team past the batch end, which still take the barriers.
// Flow runs, leader: waits until every neighbour committed inside the run
// (dependency bit set) has published its flag, then acquires at device scope
// so that the team's state reads see those commits.
__device__ __forceinline__ void flow_wait(const FlowArgs& f, const BatchArgs& a, uint32_t item, int x, int y,
                                          uint32_t kf, const WaveLds* L, int lane) {
    const uint64_t dep = f.dep[item - a.pre_base] & (kf >= 64u ? ~0ull : ((1ull << kf) - 1ull));
    if (dep == 0) return;
    uint32_t q = 0;
    bool need = (dep >> lane) & 1ull;
    if (need) {
        const int32_t o = L->nd[lane];
        q = (uint32_t)(y + packed_dy(o)) * a.W + (uint32_t)(x + packed_dx(o));
    }
    uint32_t polls = 0;
    for (;;) {
        if (need) need = __hip_atomic_load(f.flag + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != f.epoch;
        if (__ballot(need) == 0ull) break;
        ++polls;
        if ((polls & 63u) == 0u) {
            // another wave gave up: end this wait too
            const uint32_t ov = __hip_atomic_load(f.overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (ov & 4u) break;
        }
        if (polls >= f.spin_limit) {
            if (lane == 0) atomicOr(f.overflow, 4u);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Returns the result in the leader's lane 0 (zero elsewhere).
template <bool GUIDED, int FIRST_SH = 6, int KNN_U = 4>
__device__ __forceinline__ uint4 resolve_team(const BatchArgs& a, const MapDesc* maps, uint32_t item, uint32_t item0, uint32_t n_res,
                              WaveLds* Lt, int member, int T, bool valid, const float* lut, const float* glut,
                              int lane, const FlowArgs* flow = nullptr) {
    uint4 res = make_uint4(0, 0, 0, 0);
    WaveLds* L = Lt + member;
    const int slot = lane % lut_copies<GUIDED>();
    const uint32_t rel = item - item0;
    const bool lead = member == 0;
    uint32_t flat = 0;
    int x = 0, y = 0;
    uint32_t n_c = 0, n_k0 = 0, m_r = 0;  // candidates, host-list coherence entries, random ones
    uint32_t r0 = 0, nrs = 0;             // this member's random slice
    uint2 rpre = make_uint2(0, 0);
    if (valid) {
        if (lead) phase_clock(a, item, 0, lane);
        flat = gld(a.work + item);
        x = (int)(flat % a.W);
        y = (int)(flat / a.W);
        if (a.ext_cands) {
            n_c = a.ext_cand_count[rel];
            bool k0 = false;
            if (lane < (int)n_c) k0 = (a.ext_cands[(size_t)rel * a.ext_cmax + lane].y >> 31) == 0;
            n_k0 = (uint32_t)__popcll(__ballot(k0));
            m_r = n_c - n_k0;
        } else {
            m_r = a.m;
        }
        const uint32_t per = (m_r + (uint32_t)T - 1) / (uint32_t)T;
        r0 = min(m_r, (uint32_t)member * per);
        nrs = min(m_r, r0 + per) - r0;
        // the member's first 64 random candidates, loaded ahead of the kNN
        if (!a.ext_cands && !a.knn_only && (uint32_t)lane < nrs)
            rpre = gld(a.rcands + (size_t)(item - a.rc_base) * a.m + r0 + (uint32_t)lane);
    }
    // ---- 1. k nearest resolved neighbours (leader) ----
    // With a neighbour list computed ahead every member reads it and derives
    // the weights at once (in a flow run, while the leader waits for the
    // neighbours committed inside the run).
    const bool pre = valid && !a.ext_nbrs && !a.knn_only && a.pre_nbrs && item >= a.pre_base && item < a.pre_end;
    if (valid && lead) {
        const uint32_t kf0 = item_knn<KNN_U>(a, item, rel, x, y, n_res, L, lane);
        if (lane == 0) {
            L->t_kf = kf0;
            L->t_bound = ~0ull;
        }
        if (pre && kf0) item_geometry(a, x, y, kf0, L, lane);
        if (flow) flow_wait(*flow, a, item, x, y, kf0, L, lane);
    } else if (pre) {
        const uint32_t kf0 = gld(a.pre_cnt + (item - a.pre_base));
        if (lane < (int)kf0) L->nd[lane] = gld(a.pre_nbrs + (size_t)(item - a.pre_base) * a.k + lane);
        wave_sync();
        if (kf0) item_geometry(a, x, y, kf0, L, lane);
    }
    if (T > 1) __syncthreads();
    else wave_sync();
    bool done = !valid || a.knn_only;
    uint32_t kf = 0;
    if (!done) {
        kf = Lt->t_kf;
        if (!lead && !pre && lane < (int)kf) L->nd[lane] = Lt->nd[lane];
        wave_sync();
        if (lead) phase_clock(a, item, 1, lane);
        if (kf == 0) {
            if (lead) res = item_at_random(a, maps, item, rel, flat, lane);
            done = true;
        }
    }
    uint32_t n_coh = 0;
    bool exact = false;
    if (!done) {
        // ---- 2. neighbour state, weights, coherence candidates (every member) ----
        uint32_t cxy, cms;
        uint64_t uniq;
        const uint4* old_state = flow ? flow->state_old : nullptr;
        const uint64_t fresh = old_state ? flow->dep[item - a.pre_base] : 0ull;
        n_coh = item_neighbours<GUIDED>(a, maps, item, rel, x, y, kf, n_k0, L, lane, pre, cxy, cms, uniq, old_state,
                                        fresh);
        if (!a.ext_cands) {
            n_k0 = n_coh;
            n_c = n_coh + a.m;
        }
        if (lead) phase_clock(a, item, 2, lane);
        exact = a.exact || a.ext_full_cost || a.lut_nan || !(L->avg > 0.0);
        if (!exact) {
            // ---- 3. this member's slice, exact pruning ----
            // slice: unique coherence candidates [c0, c0 + ncs) then random ones
            // [r0, r0 + nrs): the coherence group alone first (it usually holds
            // the winner and bounds the rest), then the random ones 64 at a time.
            const uint32_t n_u = (uint32_t)__popcll(uniq);
            const uint32_t cper = (n_u + (uint32_t)T - 1) / (uint32_t)T;
            const uint32_t c0 = min(n_u, (uint32_t)member * cper);
            const uint32_t ncs = min(n_u, c0 + cper) - c0;
            const uint32_t nv = ncs + nrs;
            const uint32_t upos = bits_below(uniq);
            const bool mine_u = ((uniq >> lane) & 1ull) && upos >= c0 && upos < c0 + ncs;
            Best bst{FLT_MAX, kNone, 0u, 0u};
            uint32_t g_lo = 0;
            bool first = true;
            while (g_lo < nv) {
                // groups: the first kLead coherence candidates (a full evaluation of
                // a few likely winners gives an early bound), the other coherence
                // candidates, then the random ones 64 at a time
                const uint32_t g_hi = g_lo < ncs ? (g_lo == 0 && ncs > 2 * kLead ? kLead : ncs)
                                                 : min(nv, g_lo + (uint32_t)kWave);
                // coherence entries of the group
                if (mine_u) {
                    const uint32_t v = upos - c0;
                    if (v >= g_lo && v < g_hi) {
                        L->act_xy[v - g_lo] = cxy;
                        L->act_ms[v - g_lo] = cms;
                        L->act_idx[v - g_lo] = (uint32_t)lane;
                    }
                }
                // random entries of the group
                const uint32_t v = g_lo + (uint32_t)lane;
                const uint32_t rr = v >= ncs ? v - ncs : 0u;
                const uint32_t src = rr < (uint32_t)kWave ? rr : 0u;
                const uint32_t pxy = (uint32_t)__shfl((int)rpre.x, (int)src);
                const uint32_t pms = (uint32_t)__shfl((int)rpre.y, (int)src);
                if (v < g_hi && v >= ncs) {
                    const uint32_t ci = n_k0 + r0 + rr;
                    uint32_t rxy, rms;
                    if (!a.ext_cands && rr < (uint32_t)kWave) {
                        rxy = pxy;
                        rms = pms | 0x80000000u;
                    } else {
                        cand_get(a, L, rel, item, n_coh, ci, rxy, rms);
                    }
                    L->act_xy[lane] = rxy;
                    L->act_ms[lane] = rms;
                    L->act_idx[lane] = ci;
                }
                wave_sync();
                eval_group<GUIDED, FIRST_SH>(maps, a.n_maps, L, lut, glut, slot, lane, kf, g_hi - g_lo, bst,
                                             T > 1 ? &Lt->t_bound : nullptr);
                if (lead && first) phase_clock(a, item, 4, lane);
                first = false;
                g_lo = g_hi;
            }
            if (lead) phase_clock(a, item, 5, lane);
            if (lane == 0) {
                L->t_B = bst.B;
                L->t_Bi = bst.Bi;
                L->t_xy = bst.xy;
                L->t_ms = bst.ms;
            }
        } else if (lead) {
            float best = FLT_MAX;
            uint32_t bidx = 0;
            item_exact<GUIDED>(a, maps, item, rel, kf, n_coh, n_c, L, lut, glut, slot, lane, best, bidx);
            res = item_winner(a, maps, item, rel, n_coh, L, best, bidx, lane);
        }
    }
    if (T > 1) __syncthreads();
    else wave_sync();
    // ---- 4. winner: lexicographic (cost, index) minimum over the members ----
    if (!done && !exact && lead) {
        float best = FLT_MAX;
        uint32_t bidx = kNone;
        uint2 bxm = make_uint2(0, 0);
        for (int q = 0; q < T; ++q) {
            const float B = Lt[q].t_B;
            const uint32_t Bi = Lt[q].t_Bi;
            if (Bi != kNone && (bidx == kNone || B < best || (B == best && Bi < bidx))) {
                best = B;
                bidx = Bi;
                bxm = make_uint2(Lt[q].t_xy, Lt[q].t_ms);
            }
        }
        const bool have = bidx != kNone;
        if (!have) {
            best = FLT_MAX;
            bidx = 0;
        }
        res = item_winner(a, maps, item, rel, n_coh, L, best, bidx, lane, have ? &bxm : nullptr);
    }
    if (valid && lead) phase_clock(a, item, 6, lane);
    return res;
}


// Generator::update (ms.rs:334-377) for one item: colour from the stage
// level, source coordinate, ids; new pixels also enter the resolved mask.
// In tiling mode (gmask) a new pixel resolved from neighbours also gets its
// ghost copies (flush_resolved with tiling, ms.rs:296-331); one resolved at
// random does not.
__device__ __forceinline__ void commit_one(uint32_t W, uint32_t tiles_x, const MapDesc* maps, const uint32_t* work,
                                           const uint4* result, uint32_t item, uint32_t item0, uint32_t first_new,
                                           uint4* state, uint64_t* mask, float* score, uint64_t* gmask) {
    uint32_t p = gld(work + item);
    uint4 r = gld(result + (item - item0));
    uint32_t cmap = r.y & 0xffffu;
    const MapDesc md = maps[cmap];
    uint32_t sx = r.x & 0xffffu, sy = r.x >> 16;
    uint32_t color = gld(md.px + sy * md.w + sx);
    gst(state + p, make_uint4(color, r.x, r.y & ~kFallbackBit, r.z));
    if (item >= first_new) {
        uint32_t px = p % W, py = p / W;
        const uint32_t ti = (py >> 3) * tiles_x + (px >> 3);
        const uint64_t bit = 1ull << (((py & 7u) << 3) | (px & 7u));
        atomicOr((unsigned long long*)&mask[ti], bit);
        if (gmask && !(r.y & kFallbackBit)) atomicOr((unsigned long long*)&gmask[ti], bit);
        gst(score + p, __uint_as_float(r.w));
    }
}

constexpr uint32_t kMapCache = 16;  // map descriptors kept in LDS

template <bool GUIDED>
__device__ __forceinline__ WaveLds* setup_lds(const BatchArgs& a, unsigned char* smem, float*& lut, float*& glut,
                                              MapDesc* smaps, const MapDesc*& maps) {
    lut = reinterpret_cast<float*>(smem);
    constexpr int C = lut_copies<GUIDED>();
    glut = lut + 256 * C;
    for (int t = threadIdx.x; t < 256 * C; t += blockDim.x) {
        lut[t] = a.lut[t / C];
        if (GUIDED) glut[t] = a.lut[256 + t / C];
    }
    maps = a.maps;
    if (a.maps && a.n_maps <= kMapCache) {
        if (threadIdx.x < a.n_maps) smaps[threadIdx.x] = a.maps[threadIdx.x];
        maps = smaps;
    }
    int32_t* oc = offtab_lds();
    for (uint32_t t = threadIdx.x; t < kOffCache && t < a.offtab_n; t += blockDim.x) oc[t] = a.offtab[t];
    __syncthreads();
    return reinterpret_cast<WaveLds*>(smem + lut_bytes(GUIDED));
}

// Profiling: the launch's execution span on the device clock, as the
// earliest workgroup start (kept bit-inverted, so that both words take a
// maximum and zero means empty) and the latest workgroup end. Workgroups are
// dispatched in index order, so the first eight carry the earliest start;
// leaving the others out halves the atomics on the two words.
__device__ __forceinline__ void launch_span(const BatchArgs& a, uint64_t t_start) {
    if (!a.dbg_launch) return;
    __syncthreads();
    if (threadIdx.x == 0) {
        if (blockIdx.x < 8) atomicMax(&a.dbg_launch[0], ~(unsigned long long)t_start);
        atomicMax(&a.dbg_launch[1], (unsigned long long)wall_clock64());
    }
}

// One batch, teams of waves per item. With a.fused_commit the workgroup that
// finishes last commits the whole batch (the others have written their
// results), saving the commit launch.
template <bool GUIDED>
__global__ __launch_bounds__(kWave* kWavesPerWG) void k_resolve(BatchArgs a) {
    const uint64_t t_start = a.dbg_launch ? wall_clock64() : 0;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t is_last;
    __shared__ MapDesc smaps[kMapCache];
    float *lut, *glut;
    const MapDesc* maps;
    WaveLds* wls = setup_lds<GUIDED>(a, smem, lut, glut, smaps, maps);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));  // scalar LDS base
    const int lane = threadIdx.x % kWave;
    const int T = a.team > 1 ? (int)a.team : 1;
    const int team = wave / T, member = wave % T;
    const uint32_t item = a.item0 + blockIdx.x * (uint32_t)(kWavesPerWG / T) + (uint32_t)team;
    const uint4 r =
        resolve_team<GUIDED>(a, maps, item, a.item0, a.n_res, wls + team * T, member, T, item < a.item1, lut, glut, lane);
    if (a.state2) {
        if (member == 0 && lane == 0 && item < a.item1) {
            // the record goes to state2 (Generator::update, ms.rs:334-377);
            // the score is read only after the run
            const uint32_t p = a.work[item];
            const MapDesc& md = maps[r.y & 0xffffu];
            const uint32_t color = gld(md.px + (r.x >> 16) * md.w + (r.x & 0xffffu));
            a.state2[p] = make_uint4(color, r.x, r.y & ~kFallbackBit, r.z);
            a.stamp[p] = a.serial;
            if (item >= a.first_new) {
                const uint32_t px = p % a.W, py = p / a.W;
                atomicOr((unsigned long long*)&a.pend_w[(py >> 3) * a.tiles_x + (px >> 3)],
                         1ull << (((py & 7u) << 3) | (px & 7u)));
                a.score_w[p] = __uint_as_float(r.w);
            }
        }
        // the previous batch's commit, and the pending bits of the one before
        const uint32_t stride = gridDim.x * blockDim.x, t0 = blockIdx.x * blockDim.x + threadIdx.x;
        for (uint32_t i = a.prev0 + t0; i < a.prev1; i += stride) {
            const uint32_t p = a.work[i];
            a.state_w[p] = a.state2[p];
            if (i >= a.first_new) {
                const uint32_t px = p % a.W, py = p / a.W;
                atomicOr((unsigned long long*)&a.mask_w[(py >> 3) * a.tiles_x + (px >> 3)],
                         1ull << (((py & 7u) << 3) | (px & 7u)));
            }
        }
        for (uint32_t i = a.pp0 + t0; i < a.pp1; i += stride) {
            const uint32_t p = a.work[i];
            a.pend_c[((p / a.W) >> 3) * a.tiles_x + ((p % a.W) >> 3)] = 0ull;
        }
        launch_span(a, t_start);
        return;
    }
    if (!a.fused_commit) {
        launch_span(a, t_start);
        return;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) is_last = atomicAdd(a.done_counter, 1u) == gridDim.x - 1;
    __syncthreads();
    if (is_last) {
        __threadfence();
        for (uint32_t i = a.item0 + threadIdx.x; i < a.item1; i += blockDim.x)
            commit_one(a.W, a.tiles_x, maps, a.work, a.result, i, a.item0, a.first_new, a.state_w, a.mask_w,
                       a.score_w, a.gmask);
        if (threadIdx.x == 0) *a.done_counter = 0u;
    }
    launch_span(a, t_start);
}

// One batch of single-wave items on persistent workgroups (at most as many
// as are resident): workgroup w takes an equal slice of the batch, and its
// waves take the slice's items from an LDS counter until they run out, so a
// wave never waits for the slowest item of its workgroup, and the cost
// tables are set up once per workgroup. (One device-wide counter for every
// wave serialised the item handout.) Results go to a.result as in k_resolve.
template <bool GUIDED>
__global__ __launch_bounds__(kWave* kWavesPerWG) __attribute__((amdgpu_waves_per_eu(4))) void k_resolve_persist(
    BatchArgs a) {
    const uint64_t t_start = a.dbg_launch ? wall_clock64() : 0;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ MapDesc smaps[kMapCache];
    __shared__ uint32_t s_next;
    if (threadIdx.x == 0) s_next = 0;
    float *lut, *glut;
    const MapDesc* maps;
    WaveLds* wls = setup_lds<GUIDED>(a, smem, lut, glut, smaps, maps);  // ends with a barrier
    // wave-uniform in a scalar register: the wave's LDS block is then a
    // scalar base, not a vector address held across the item loop
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int lane = threadIdx.x % kWave;
    const uint32_t n = a.item1 - a.item0;
    const uint32_t s0 = (uint32_t)((uint64_t)blockIdx.x * n / gridDim.x);
    const uint32_t s1 = (uint32_t)((uint64_t)(blockIdx.x + 1) * n / gridDim.x);
    // the launcher admits a.team <= 1 only; kept a runtime value, as in k_resolve
    const int T = a.team > 1 ? (int)a.team : 1;
    const int member = wave % T;
    for (;;) {
        uint32_t idx = 0;
        if (lane == 0) idx = s0 + atomicAdd(&s_next, 1u);
        idx = (uint32_t)__shfl((int)idx, 0);
        if (idx >= s1) break;
        // an opaque lane index: the item's lane-dependent addresses are
        // computed per item instead of held in registers across the loop
        int ln = lane;
        asm volatile("" : "+v"(ln));
        resolve_team<GUIDED>(a, maps, a.item0 + idx, a.item0, a.n_res, wls + (wave - member), member, T, true, lut,
                             glut, ln);
    }
    launch_span(a, t_start);
}

// A run of small consecutive batches (at most kSmallWaves items each) in one
// workgroup: resolve, barrier, commit, barrier, next batch. Batch b covers
// items [bounds[b], bounds[b+1]) with nres[b] resolved pixels at its start.
template <bool GUIDED>
__global__ __launch_bounds__(kWave* kSmallWaves) void k_resolve_small(BatchArgs a, const uint32_t* bounds,
                                                                      const uint32_t* nres, uint32_t n_batches) {
    const uint64_t t_start = a.dbg_launch ? wall_clock64() : 0;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ MapDesc smaps[kMapCache];
    float *lut, *glut;
    const MapDesc* maps;
    WaveLds* wls = setup_lds<GUIDED>(a, smem, lut, glut, smaps, maps);
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    for (uint32_t b = 0; b < n_batches; ++b) {
        const uint32_t i0 = bounds[b], i1 = bounds[b + 1];
        // team size: the largest power of two with one team per item
        int T = kSmallWaves;
        while (T > 1 && (uint32_t)T * (i1 - i0) > (uint32_t)kSmallWaves) T >>= 1;
        const int team = wave / T, member = wave % T;
        const uint32_t item = i0 + (uint32_t)team;
        resolve_team<GUIDED>(a, maps, item, i0, nres[b], wls + team * T, member, T, item < i1, lut, glut, lane);
        __syncthreads();
        const uint32_t ci = i0 + threadIdx.x;
        if (ci < i1) commit_one(a.W, a.tiles_x, maps, a.work, a.result, ci, i0, a.first_new, a.state_w, a.mask_w,
                                a.score_w, a.gmask);
        __syncthreads();
    }
    launch_span(a, t_start);
}

// A run of consecutive latency-bound batches in one cooperative launch (all
// workgroups resident): per batch, resolve (teams of waves sized to the
// grid), grid barrier, commit spread over the grid, grid barrier. Every
// workgroup takes the same 2 * n_batches barriers. A batch larger than the
// grid holds (one item per wave) sets *overflow and is skipped; the host
// never forms one.
template <bool GUIDED>
__global__ __launch_bounds__(kWave* kSmallWaves) void k_resolve_coop(BatchArgs a, const uint32_t* bounds,
                                                                     const uint32_t* nres, uint32_t n_batches,
                                                                     uint32_t team_max, uint32_t* overflow) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ MapDesc smaps[kMapCache];
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    float *lut, *glut;
    const MapDesc* maps;
    WaveLds* wls = setup_lds<GUIDED>(a, smem, lut, glut, smaps, maps);
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const uint32_t waves = gridDim.x * (uint32_t)kSmallWaves;
    for (uint32_t b = 0; b < n_batches; ++b) {
        const uint32_t i0 = bounds[b], i1 = bounds[b + 1], n = i1 - i0;
        if (n <= waves) {
            int T = (int)min(team_max, (uint32_t)kSmallWaves);
            while (T > 1 && (uint32_t)T * n > waves) T >>= 1;
            const int team = wave / T, member = wave % T;
            const uint32_t item = i0 + blockIdx.x * (uint32_t)(kSmallWaves / T) + (uint32_t)team;
            resolve_team<GUIDED>(a, maps, item, i0, nres[b], wls + team * T, member, T, item < i1, lut, glut, lane);
        } else if (blockIdx.x == 0 && threadIdx.x == 0) {
            atomicOr(overflow, 2u);
        }
        grid.sync();
        if (n <= waves)
            for (uint32_t ci = i0 + blockIdx.x * blockDim.x + threadIdx.x; ci < i1; ci += gridDim.x * blockDim.x)
                commit_one(a.W, a.tiles_x, maps, a.work, a.result, ci, i0, a.first_new, a.state_w, a.mask_w,
                           a.score_w, a.gmask);
        grid.sync();
    }
}

// A dependency-ordered run of new-pixel items (FlowArgs): one item per
// workgroup at a time, the workgroup's waves as its team. Thread 0 is the
// team leader's lane 0: it holds the item's result, commits the pixel and
// publishes it with a device-scope release. a.result is not used.
template <bool GUIDED>
__global__ __launch_bounds__(kWave* kWavesPerWG) void k_resolve_flow(BatchArgs a, FlowArgs f) {
    const uint64_t t_start = a.dbg_launch ? wall_clock64() : 0;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ MapDesc smaps[kMapCache];
    __shared__ uint32_t s_idx;
    float *lut, *glut;
    const MapDesc* maps;
    WaveLds* wls = setup_lds<GUIDED>(a, smem, lut, glut, smaps, maps);
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int T = (int)(blockDim.x / kWave);
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_idx = atomicAdd(f.ticket, 1u) - f.ticket_base;
        __syncthreads();
        const uint32_t idx = s_idx;
        if (idx >= f.n) break;
        const uint32_t item = a.item0 + idx;
        const uint32_t p = threadIdx.x == 0 ? a.work[item] : 0u;
        const uint4 r = resolve_team<GUIDED>(a, maps, item, item, 0u, wls, wave, T, true, lut, glut, lane, &f);
        if (threadIdx.x == 0) {
            // Generator::update (ms.rs:334-377): the state record, published
            // at once; the mask bit and the score are read only after the run
            const MapDesc& md = maps[r.y & 0xffffu];
            const uint32_t color = gld(md.px + (r.x >> 16) * md.w + (r.x & 0xffffu));
            a.state_w[p] = make_uint4(color, r.x, r.y & ~kFallbackBit, r.z);
            __hip_atomic_store(f.flag + p, f.epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            if (!f.state_old) {
                // a new pixel enters the resolved mask and keeps its first score
                const uint32_t px = p % a.W, py = p / a.W;
                atomicOr((unsigned long long*)&a.mask_w[(py >> 3) * a.tiles_x + (px >> 3)],
                         1ull << (((py & 7u) << 3) | (px & 7u)));
                a.score_w[p] = __uint_as_float(r.w);
            }
            phase_clock(a, item, 10, 0);
        }
    }
    launch_span(a, t_start);
}

// Neighbour lists of items [item0, item1) ahead of their batches, one
// wavefront per item (the resolve kernels read them instead of searching).
__device__ __forceinline__ void prefetch_item(const PrefetchArgs& a, WaveLds* L, uint32_t item, int lane) {
    const uint32_t rel = item - a.item0;
    const uint32_t flat = gld(a.work + item);
    uint32_t n_at, bs = 0;
    if (a.n_at) {
        n_at = a.n_at[rel];
    } else {
        uint32_t lo = 0, hi = a.nb - 1;  // last batch starting at or before the item
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (a.bounds[mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        n_at = a.bnres[lo];
        bs = a.bounds[lo];
    }
    const KnnView v{a.W,        a.H,      a.tiles_x, a.tiles_y, a.mask, a.rlist, a.offtab, a.offtab_n,
                    a.mask_new, a.tile_min, a.rank, n_at};
    uint32_t iters, buffered;
    const uint32_t kf = knn_search(v, L, (int)(flat % a.W), (int)(flat / a.W), a.k, n_at, lane, iters, buffered);
    if (lane < (int)kf) gst(a.out_nbrs + (size_t)rel * a.k + lane, L->nd[lane]);
    if (lane == 0) gst(a.out_cnt + rel, kf);
    if (a.out_dep) {
        bool d = false;
        if (lane < (int)kf) {
            const int32_t o = L->nd[lane];
            const int X = (int)(flat % a.W) + packed_dx(o), Y = (int)(flat / a.W) + packed_dy(o);
            if (item < a.redo_end) {
                // redone by an earlier batch of the segment
                const uint32_t r = a.rank[(uint32_t)Y * a.W + (uint32_t)X];
                d = r >= a.locked && r - a.locked < bs;
            } else if (a.mask_new) {
                d = px_set(a.mask_new, a.tiles_x, X, Y);
            }
        }
        const uint64_t dm = __ballot(d);
        if (lane == 0) a.out_dep[rel] = dm;
    }
}

__global__ __launch_bounds__(kWave* kWavesPerWG) void k_knn_prefetch(PrefetchArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    WaveLds* L = reinterpret_cast<WaveLds*>(smem) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const uint32_t item = a.item0 + blockIdx.x * kWavesPerWG + threadIdx.x / kWave;
    if (item >= a.item1) return;
    prefetch_item(a, L, item, lane);
}

__global__ void k_unpack(UnpackArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n) {
        const uint4 st = a.state[i];
        if (a.rgba) a.rgba[i] = st.x;
        if (a.coord) {
            a.coord[3 * i] = st.y & 0xffffu;
            a.coord[3 * i + 1] = st.y >> 16;
            a.coord[3 * i + 2] = st.z & 0xffffu;
        }
        if (a.ids) {
            a.ids[2 * i] = st.w;
            a.ids[2 * i + 1] = st.z >> 16;
        }
    }
    if (a.rscore && i < a.n_res) a.rscore[i] = a.score[a.rlist[i]];
}

// get_id_maps (ms.rs:603-630): each channel is gen_range(0..255) of its own
// seeded stream; seeds are u32 arithmetic on the ids (wrapping) widened to u64.
__device__ __forceinline__ uint32_t id_colour(uint64_t s0, uint64_t s1, uint64_t s2) {
    Pcg32 g0 = Pcg32::seed_from_u64(s0), g1 = Pcg32::seed_from_u64(s1), g2 = Pcg32::seed_from_u64(s2);
    const uint32_t r = range_small(g0, 0, 255), g = range_small(g1, 0, 255), b = range_small(g2, 0, 255);
    return r | (g << 8) | (b << 16) | 0xff000000u;
}

__global__ void k_id_maps(const uint4* state, uint32_t n, uint32_t* patch_rgba, uint32_t* map_rgba) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 st = state[i];
    const uint32_t patch = st.w, map = st.z >> 16;
    if (patch_rgba) patch_rgba[i] = id_colour(patch, patch * 5u + 21u, patch / 4u + 12u);
    if (map_rgba) map_rgba[i] = id_colour((uint64_t)map * 200u, map * 5u + 341u, map * 1200u - 35412u);
}

// get_uncertainty_map (ms.rs:633-653): n = (min(score, 1) * 255) as u8,
// colour [n, 255 - n, 0, 255], for every pixel of the resolved list.
__global__ void k_uncertainty(const uint32_t* rlist, const float* score, uint32_t n_res, uint32_t* out) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_res) return;
    const uint32_t p = rlist[q];
    const float v = fminf(score[p], 1.0f) * 255.0f;
    // Rust's float-to-u8 cast saturates (NaN to 0)
    const uint32_t c = v == v ? (uint32_t)fminf(fmaxf(v, 0.0f), 255.0f) : 0u;
    out[p] = c | ((255u - c) << 8) | 0xff000000u;
}

__global__ void k_xform_apply(const uint32_t* coord, uint32_t n, const XformSource* srcs, uint32_t n_src,
                              uint32_t* out, uint32_t* bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = coord[3 * i], y = coord[3 * i + 1], m = coord[3 * i + 2];
    if (m >= n_src || x >= srcs[m].w || y >= srcs[m].h) {
        out[i] = 0u;
        atomicOr(bad, 1u);
        return;
    }
    out[i] = srcs[m].px[(size_t)y * srcs[m].w + x];
}

__global__ void k_rank(const uint32_t* rlist, uint32_t lo, uint32_t hi, uint32_t* rank) {
    const uint32_t q = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (q < hi) rank[rlist[q]] = q;
}

__global__ void k_mask_or(const uint32_t* rlist, uint32_t lo, uint32_t hi, uint32_t W, uint32_t tiles_x,
                          uint64_t* mask, uint32_t* tile_min) {
    const uint32_t q = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= hi) return;
    const uint32_t p = rlist[q], px = p % W, py = p / W;
    const uint32_t t = (py >> 3) * tiles_x + (px >> 3);
    atomicOr((unsigned long long*)&mask[t], 1ull << (((py & 7u) << 3) | (px & 7u)));
    if (tile_min) atomicMin(&tile_min[t], q);
}

// Resolved bits of a one-byte-per-pixel mask (nonzero = resolved), one u64
// per 8x8 tile (bit = row * 8 + col): one thread per tile.
__global__ void k_mask_tiles(const uint8_t* mask, uint32_t W, uint32_t H, uint32_t tiles_x, uint32_t tiles_y,
                             uint64_t* tiles) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tiles_x * tiles_y) return;
    const uint32_t x0 = (t % tiles_x) * 8u, y0 = (t / tiles_x) * 8u;
    uint64_t bits = 0;
    for (uint32_t r = 0; r < 8u && y0 + r < H; ++r)
        for (uint32_t c = 0; c < 8u && x0 + c < W; ++c)
            if (mask[(size_t)(y0 + r) * W + x0 + c]) bits |= 1ull << ((r << 3) | c);
    tiles[t] = bits;
}

// kNN lists packed as (dx & 0xffff | dy << 16) per entry, to (x, y) int32
// pairs of the query's neighbours; entries past a query's count are (0, 0).
__global__ void k_knn_unpack(const uint32_t* q_flat, const int32_t* packed, const uint32_t* counts, uint32_t n_q,
                             uint32_t k, uint32_t W, int32_t* pairs) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_q * k) return;
    const uint32_t q = e / k, j = e % k;
    const int32_t x = (int32_t)(q_flat[q] % W), y = (int32_t)(q_flat[q] / W);
    const int32_t o = packed[e];
    const bool v = j < counts[q];
    pairs[2 * (size_t)e] = v ? x + (int32_t)(int16_t)(o & 0xffff) : 0;
    pairs[2 * (size_t)e + 1] = v ? y + (o >> 16) : 0;
}

// Random candidates of items [item0, item1): Pcg32::seed_from_u64(stage
// seed + i + 1), m draws of (map, then x and y until the sample mask accepts).
__device__ __forceinline__ void rand_item(const RandArgs& a, uint32_t item) {
    Pcg32 g = Pcg32::seed_from_u64(a.stage_seed + (uint64_t)item + 1u);
    uint2* out = a.out + (size_t)(item - a.item0) * a.m;
    for (uint32_t r = 0; r < a.m; ++r) {
        uint32_t map = (uint32_t)range_u64(g, 0, a.n_maps);
        const MapDesc md = a.maps[map];
        uint32_t x, y, tries = 0;
        for (;;) {
            x = range_u32(g, 0, md.w);
            y = range_u32(g, 0, md.h);
            if (md.smask == nullptr || md.smask[y * md.w + x] != 0) break;
            if (++tries >= (1u << 24)) {
                atomicOr(a.overflow, 1u);
                break;
            }
        }
        out[r] = make_uint2(x | (y << 16), map);
    }
}

__global__ void k_rand(RandArgs a) {
    const uint32_t item = a.item0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (item < a.item1) rand_item(a, item);
}

__global__ void k_commit(CommitArgs a) {
    uint32_t item = a.item0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= a.item1) return;
    commit_one(a.W, a.tiles_x, a.maps, a.work, a.result, item, a.item0, a.first_new, a.state, a.mask, a.score,
               a.gmask);
}

// DeferFlushArgs: commit the last deferred batch, clear the pending bits.
__global__ void k_defer_flush(DeferFlushArgs a) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = a.prev0 + t;
    if (i < a.prev1) {
        const uint32_t p = a.work[i];
        a.state[p] = a.state2[p];
        const uint32_t px = p % a.W, py = p / a.W;
        const uint32_t ti = (py >> 3) * a.tiles_x + (px >> 3);
        if (i >= a.first_new) atomicOr((unsigned long long*)&a.mask[ti], 1ull << (((py & 7u) << 3) | (px & 7u)));
        a.pend_prev[ti] = 0ull;
    }
    const uint32_t q = a.pp0 + t;
    if (q < a.pp1) {
        const uint32_t p = a.work[q];
        a.pend_pp[((p / a.W) >> 3) * a.tiles_x + ((p % a.W) >> 3)] = 0ull;
    }
}

// old[p] = state[p] for the first n pixels of the resolved list.
__global__ void k_state_snapshot(const uint32_t* rlist, uint32_t n, const uint4* state, uint4* old) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = rlist[i];
    old[p] = state[p];
}

// next_pyramid_level (ms.rs:687-700): recolour every resolved pixel.
__global__ void k_recolor(RecolorArgs a) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    uint32_t p = a.rlist[i];
    uint4 st = a.state[p];
    const MapDesc md = a.maps[st.z & 0xffffu];
    reinterpret_cast<uint32_t*>(&a.state[p])[0] = md.px[(st.y >> 16) * md.w + (st.y & 0xffffu)];
}

// image 0.23 vertical_sample: u8 in, f32 out, taps summed in order.
__global__ void k_resample_v(const uint32_t* src, uint32_t w, uint32_t h, float4* dst, uint32_t nh,
                             const uint32_t* left, const uint32_t* cnt, const float* ws, uint32_t stride) {
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t oy = blockIdx.y;
    if (x >= w || oy >= nh) return;
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;
    uint32_t l = left[oy], n = cnt[oy];
    for (uint32_t i = 0; i < n; ++i) {
        float wt = ws[oy * stride + i];
        uint32_t p = src[(size_t)(l + i) * w + x];
        t0 = __fadd_rn(t0, __fmul_rn((float)(p & 255u), wt));
        t1 = __fadd_rn(t1, __fmul_rn((float)((p >> 8) & 255u), wt));
        t2 = __fadd_rn(t2, __fmul_rn((float)((p >> 16) & 255u), wt));
        t3 = __fadd_rn(t3, __fmul_rn((float)(p >> 24), wt));
    }
    dst[(size_t)oy * w + x] = make_float4(t0, t1, t2, t3);
}

__device__ __forceinline__ uint32_t to_u8(float t) {
    float v = t < 0.0f ? 0.0f : (t > 255.0f ? 255.0f : t);
    return (uint32_t)roundf(v);
}

// image 0.23 horizontal_sample: f32 in, clamped and rounded u8 out.
__global__ void k_resample_h(const float4* src, uint32_t w, uint32_t h, uint32_t* dst, uint32_t nw,
                             const uint32_t* left, const uint32_t* cnt, const float* ws, uint32_t stride) {
    uint32_t ox = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t y = blockIdx.y;
    if (ox >= nw || y >= h) return;
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;
    uint32_t l = left[ox], n = cnt[ox];
    for (uint32_t i = 0; i < n; ++i) {
        float wt = ws[ox * stride + i];
        float4 p = src[(size_t)y * w + l + i];
        t0 = __fadd_rn(t0, __fmul_rn(p.x, wt));
        t1 = __fadd_rn(t1, __fmul_rn(p.y, wt));
        t2 = __fadd_rn(t2, __fmul_rn(p.z, wt));
        t3 = __fadd_rn(t3, __fmul_rn(p.w, wt));
    }
    dst[(size_t)y * nw + ox] = to_u8(t0) | (to_u8(t1) << 8) | (to_u8(t2) << 16) | (to_u8(t3) << 24);
}

__global__ void k_stats(StatsArgs a) {
    __shared__ double ssum[256];
    __shared__ unsigned long long scnt[256];
    double s = 0.0;
    unsigned long long c = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        float v = a.item_score[i];
        if (v >= 0.0f) {
            s += (double)v;
            c += 1;
        }
    }
    ssum[threadIdx.x] = s;
    scnt[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            ssum[threadIdx.x] += ssum[threadIdx.x + o];
            scnt[threadIdx.x] += scnt[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(a.sum, ssum[0]);
        atomicAdd(a.count, scnt[0]);
    }
}

// Sum and count of the non-negative scores of stage items [lo, hi) (at most
// kStatsChunk), in a fixed order: threads 0..255 take items t, t + 256, ...
// in turn, then a tree over the threads. Every thread of the block calls it
// (threads from 256 on only take the barriers); red / redn: 256 entries of
// LDS each. The result is valid in thread 0.
__device__ __forceinline__ void stats_chunk(const float* score, uint32_t lo, uint32_t hi, double* red,
                                            unsigned long long* redn, double& sum, unsigned long long& cnt) {
    const uint32_t t = threadIdx.x;
    if (t < 256) {
        double s = 0.0;
        unsigned long long c = 0;
        for (uint32_t i = lo + t; i < hi; i += 256) {
            const float v = score[i];
            if (v >= 0.0f) {
                s += (double)v;
                c += 1;
            }
        }
        red[t] = s;
        redn[t] = c;
    }
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (t < o) {
            red[t] += red[t + o];
            redn[t] += redn[t + o];
        }
        __syncthreads();
    }
    sum = red[0];
    cnt = redn[0];
}

__global__ __launch_bounds__(256) void k_stats_part(const float* score, uint32_t n, double* part,
                                                    unsigned long long* part_n) {
    __shared__ double red[256];
    __shared__ unsigned long long redn[256];
    const uint32_t lo = blockIdx.x * kStatsChunk;
    const uint32_t hi = min(n, lo + kStatsChunk);
    double s;
    unsigned long long c;
    stats_chunk(score, lo, hi, red, redn, s, c);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s;
        part_n[blockIdx.x] = c;
    }
}

__device__ __forceinline__ void stats_final(const double* part, const unsigned long long* part_n, uint32_t n_parts,
                                            double* sum, unsigned long long* count) {
    double s = 0.0;
    unsigned long long c = 0;
    for (uint32_t u = 0; u < n_parts; ++u) {
        s += part[u];
        c += part_n[u];
    }
    *sum = s;
    *count = c;
}

__global__ void k_stats_final(const double* part, const unsigned long long* part_n, uint32_t n_parts, double* sum,
                              unsigned long long* count) {
    if (threadIdx.x == 0 && blockIdx.x == 0) stats_final(part, part_n, n_parts, sum, count);
}

// ---- work-queue engine ----

// XCD (XCC) this wave runs on: hardware register XCC_ID, bits [3:0].
__device__ __forceinline__ uint32_t xcc_id() {
    return (uint32_t)__builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 7u;
}

// Acquire: following loads see what other workgroups of the XCD released
// (device-scope acquire: the CU's L1 and the L2's non-coherent lines are
// dropped).
__device__ __forceinline__ void unit_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// Release to the workgroups of the same XCD: the wave waits until its stores
// have completed to the XCD's L2, which every CU of the XCD reads through.
// A job's workgroups all run on its XCD, so the L2 write-back of a
// device-scope release is not needed (it made the engine four times slower).
__device__ __forceinline__ void unit_release() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// A spread launch's jobs have workgroups on every XCD: after the workgroup's
// waves have released to the L2 and met at the barrier, one thread writes
// the L2 back (device-scope release) before the unit is counted finished.
__device__ __forceinline__ void unit_publish(bool spread) {
    if (spread) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}

// Weight-exp sweep (test hook): out[i - i0] = (float)exp_dd(-1.0 * x_i), the
// expression of the neighbour weights in item_geometry, for the arguments
// x_i = t * 8 when i % 4 != 0, else 8 + t * 56, with t = ((i * 0x9E3779B97F4A7C15)
// >> 11) * 2^-53 in [0, 1): three quarters of the arguments in [0, 8), where
// d / mean lies for real neighbour sets, the rest up to 64.
__global__ void k_exp_sweep(uint64_t i0, uint64_t n, float* out) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint64_t i = i0 + r;
    const double t = (double)((i * 0x9E3779B97F4A7C15ull) >> 11) * 0x1p-53;
    const double x = (i & 3u) ? t * 8.0 : __dadd_rn(8.0, __dmul_rn(t, 56.0));
    out[r] = (float)exp_dd(-1.0 * x);
}

__global__ void k_xcc_probe(uint32_t* out, uint32_t n) {
    if (threadIdx.x == 0 && blockIdx.x < n) out[blockIdx.x] = xcc_id();
}

// Persistent workgroups (as many as are resident) that drain the programs of
// the jobs placed on their XCD. Wave 0 finds a unit: its lanes read up to 64
// jobs' cursors at a time, from the last job served on; a job whose current
// phase still has units left is claimed with an atomic add on its cursor
// (a claim that lands past the phase's units, because others took them in
// between, is dropped and the search goes on). Then every wave works on the
// unit, the workgroup counts it finished, and the workgroup that finishes a
// phase's last unit closes it: a fused batch commit or the stage statistics'
// final sum, then the next phase's cursor. A workgroup stops when every job of
// its XCD is through its program, or after idle_ticks without any unit
// finishing on its XCD (setting its XCD's error word).
//
// Ordering: a unit's stores complete to the L2 before its workgroup counts
// the unit finished; a workgroup acquires at device scope when it takes a
// unit of another phase than its last one, and before closing a phase. All of
// a job's workgroups run on its XCD, so the L2 they share carries a phase's
// results to the next phase. (Dropping only the CU's L1 at the acquire was
// measured not to be enough: with many workgroups on one job, results then
// differed from the job run alone.)
//
// Occupancy: bounded to 6 waves per SIMD (80 VGPRs, three workgroups per CU)
// with some 45 VGPRs spilled, against 4 waves (2 workgroups) without spills:
// the 256-job mix ran 1.17 s against 1.28 s.
template <bool GUIDED>
__global__ __launch_bounds__(kWave* kWavesPerWG) __attribute__((amdgpu_waves_per_eu(6))) void k_engine(QEngineArgs e, const BatchArgs* __restrict__ args) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ MapDesc smaps[kMapCache];
    __shared__ uint32_t s_job, s_ph, s_unit, s_stop, s_last, s_key;
    constexpr int C = lut_copies<GUIDED>();
    float* lut = reinterpret_cast<float*>(smem);
    float* glut = lut + 256 * C;
    WaveLds* wls = reinterpret_cast<WaveLds*>(smem + lut_bytes(GUIDED));
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int lane = threadIdx.x % kWave;
    // the search's own state lives in LDS: nothing of it is held in registers
    // across a unit's work
    __shared__ uint32_t s_hint, s_seen, s_last_job, s_last_ph;
    __shared__ unsigned long long s_idle_since;
    int32_t* oc = offtab_lds();
    for (uint32_t t = threadIdx.x; t < kOffCache && t < e.offtab_n; t += blockDim.x) oc[t] = e.offtab[t];
    // the control words this workgroup uses: its XCD's, or XCD 0's for a spread launch
    const bool spread = (e.flags & kEngineSpread) != 0u;
    const uint32_t home = spread ? 0u : xcc_id();
    if (threadIdx.x == 0) {
        const uint32_t x = home;
        const uint32_t j0 = e.ctl->job0[x], nj = e.ctl->job0[x + 1] - j0;
        s_key = 0;
        s_hint = nj ? j0 + (blockIdx.x / 8u) % nj : j0;
        s_idle_since = 0;
        s_seen = 0;
        s_last_job = s_last_ph = ~0u;
    }
    __shared__ unsigned long long s_t[4];
    __shared__ uint32_t s_nu[2];
    if (threadIdx.x < 4) s_t[threadIdx.x] = 0;
    if (threadIdx.x < 2) s_nu[threadIdx.x] = 0;
    for (;;) {
        __syncthreads();
        const uint64_t c0 = wall_clock64();  // scalar: no VGPRs held across the unit
        if (wave == 0) {
            const uint32_t xcc = home;
            const uint32_t j0 = e.ctl->job0[xcc], nj = e.ctl->job0[xcc + 1] - j0;
            const uint32_t hint = s_hint;
            uint64_t idle_since = s_idle_since;
            uint32_t seen = s_seen;
            uint32_t gj = ~0u, gph = 0, gu = 0, stop = 0;
            for (;;) {
                bool any = false;
                for (uint32_t base = 0; base < nj && gj == ~0u; base += kWave) {
                    const uint32_t r = base + (uint32_t)lane;
                    uint32_t jj = 0;
                    bool avail = false;
                    if (r < nj) {
                        jj = j0 + (hint - j0 + r) % nj;
                        const unsigned long long c =
                            __hip_atomic_load(&e.dyn[jj].cursor, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        const uint32_t ph = (uint32_t)(c >> 32);
                        if (ph < e.jobs[jj].n_phases) avail = (uint32_t)c < e.jobs[jj].phases[ph].units;
                    }
                    uint64_t m = __ballot(avail);
                    while (m) {
                        any = true;
                        const int src = __ffsll((unsigned long long)m) - 1;
                        m &= m - 1;
                        const uint32_t cj = (uint32_t)__shfl((int)jj, src);
                        unsigned long long old = 0;
                        if (lane == 0)
                            old = __hip_atomic_fetch_add(&e.dyn[cj].cursor, 1ull, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
                        const uint32_t ph = (uint32_t)__shfl((int)(uint32_t)(old >> 32), 0);
                        const uint32_t u = (uint32_t)__shfl((int)(uint32_t)old, 0);
                        if (ph < e.jobs[cj].n_phases && u < e.jobs[cj].phases[ph].units) {
                            gj = cj;
                            gph = ph;
                            gu = u;
                            break;
                        }
                    }
                }
                if (gj != ~0u) break;
                // nothing to take: done when every job of the XCD is through
                if (__hip_atomic_load(&e.ctl->x[xcc].active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
                    stop = 1;
                    break;
                }
                const uint32_t pg = __hip_atomic_load(&e.ctl->x[xcc].progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t now = wall_clock64();
                if (idle_since == 0 || pg != seen) {
                    idle_since = now;
                    seen = pg;
                } else if (now - idle_since > e.idle_ticks) {
                    if (lane == 0) atomicOr(&e.ctl->x[xcc].error, 1u);
                    stop = 1;
                    break;
                }
                if (!any) __builtin_amdgcn_s_sleep(8);
            }
            const bool same = gj == s_last_job && gph == s_last_ph;
            if (lane == 0) {
                s_job = gj;
                s_ph = gph;
                s_unit = gu;
                s_stop = stop;
                s_seen = seen;
                s_idle_since = stop ? idle_since : 0;
                if (!stop) {
                    s_hint = gj;
                    s_last_job = gj;
                    s_last_ph = gph;
                }
            }
            // a workgroup that stays in one phase has seen its inputs already
            if (!stop && !same) unit_acquire();
        }
        __syncthreads();
        const uint64_t c1 = wall_clock64();
        if (threadIdx.x == 0) s_t[0] += c1 - c0;
        if (s_stop) break;
        // uniform: the job's and the phase's arguments go to scalar registers
        const uint32_t jb = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_job);
        const uint32_t phi = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_ph);
        const uint32_t unit = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_unit);
        const QJob J = ld_uniform(e.jobs + jb);
        const QPhase P = ld_uniform(J.phases + phi);
        const uint32_t tid = threadIdx.x;
        if (P.kind == Q_RESOLVE) {
            BatchArgs a = ld_uniform(args + P.arg);
            a.item0 = P.i0;
            a.item1 = P.i1;
            a.n_res = P.n_res;
            a.team = P.team;
            a.rc_base = P.rc_base;
            const bool pre = P.pre_end > P.pre_base;
            a.pre_nbrs = pre ? J.pre_nbrs : nullptr;
            a.pre_cnt = pre ? J.pre_cnt : nullptr;
            a.pre_base = P.pre_base;
            a.pre_end = P.pre_end;
            // the job stage's cost tables and map descriptors in LDS
            const uint32_t key = ((jb << 4) | P.stage) + 1u;
            const bool reload = s_key != key;
            if (reload) {
                for (int t = (int)tid; t < 256 * C; t += (int)blockDim.x) {
                    lut[t] = a.lut[t / C];
                    if (GUIDED) glut[t] = a.lut[256 + t / C];
                }
                if (a.n_maps <= kMapCache && tid < a.n_maps) smaps[tid] = a.maps[tid];
                __syncthreads();
            }
            const MapDesc* maps = a.n_maps <= kMapCache ? smaps : a.maps;
            const int T = (int)P.team;
            const int team = wave / T, member = wave % T;
            const uint32_t per = (uint32_t)(kWavesPerWG / T);
            for (uint32_t r = 0; r < P.rep; ++r) {
                const uint32_t item = P.i0 + (unit * P.rep + r) * per + (uint32_t)team;
                int ln = lane;
                asm volatile("" : "+v"(ln));
                resolve_team<GUIDED, kEngineFirstChunk, kEngineKnnU>(a, maps, item, a.item0, a.n_res, wls + team * T,
                                                                     member, T, item < a.item1, lut, glut, ln);
            }
            if (reload && tid == 0) s_key = key;
        } else if (P.kind == Q_COMMIT) {
            const BatchArgs a = ld_uniform(args + P.arg);
            const uint32_t i = P.i0 + unit * kQElems + tid;
            if (tid < kQElems && i < P.i1)
                commit_one(a.W, a.tiles_x, a.maps, a.work, J.result, i, P.i0, a.first_new, a.state_w, a.mask_w,
                           a.score_w, a.gmask);
        } else if (P.kind == Q_RAND) {
            const BatchArgs a = ld_uniform(args + P.arg);
            const RandArgs ra{a.maps, a.n_maps, a.stage_seed, P.i0, P.i1, a.m, J.rc, J.overflow};
            const uint32_t i = P.i0 + unit * kQElems + tid;
            if (tid < kQElems && i < P.i1) rand_item(ra, i);
        } else if (P.kind == Q_RECOLOR) {
            const BatchArgs a = ld_uniform(args + P.arg);
            const uint32_t i = P.i0 + unit * kQElems + tid;
            if (tid < kQElems && i < P.i1) {
                const uint32_t p = a.rlist[i];
                const uint4 st = a.state_w[p];
                const MapDesc md = a.maps[st.z & 0xffffu];
                reinterpret_cast<uint32_t*>(&a.state_w[p])[0] = md.px[(st.y >> 16) * md.w + (st.y & 0xffffu)];
            }
        } else if (P.kind == Q_STATS) {
            const BatchArgs a = ld_uniform(args + P.arg);
            const uint32_t lo = P.i0 + unit * kStatsChunk;
            const uint32_t hi = min(P.i1, lo + kStatsChunk);
            double* red = reinterpret_cast<double*>(wls);
            unsigned long long* redn = reinterpret_cast<unsigned long long*>(red + 256);
            double sm;
            unsigned long long cn;
            stats_chunk(a.item_score, lo, hi, red, redn, sm, cn);
            if (tid == 0) {
                J.st_part[unit] = sm;
                J.st_part_n[unit] = cn;
            }
        } else if (P.kind == Q_CLEAR) {
            const uint32_t t = P.i0 + unit * kQElems + tid;
            if (tid < kQElems && t < P.i1) {
                J.mend[t] = 0ull;
                J.tmin[t] = 0xffffffffu;
            }
        } else if (P.kind == Q_MASKOR) {
            const BatchArgs a = ld_uniform(args + P.arg);
            const uint32_t q = P.i0 + unit * kQElems + tid;
            if (tid < kQElems && q < P.i1) {
                const uint32_t p = a.rlist[q], px = p % a.W, py = p / a.W;
                const uint32_t t = (py >> 3) * a.tiles_x + (px >> 3);
                atomicOr((unsigned long long*)&J.mend[t], 1ull << (((py & 7u) << 3) | (px & 7u)));
                atomicMin(&J.tmin[t], q);
            }
        } else if (P.kind == Q_PREFETCH) {
            const BatchArgs a = ld_uniform(args + P.arg);
            PrefetchArgs pa{};
            pa.W = a.W;
            pa.H = a.H;
            pa.tiles_x = a.tiles_x;
            pa.tiles_y = a.tiles_y;
            pa.mask = a.mask;
            if (P.pre_end) {
                pa.mask_new = J.mend;
                pa.tile_min = J.tmin;
                pa.rank = J.rank;
            }
            pa.rlist = a.rlist;
            pa.work = a.work;
            pa.n_at = J.n_at + P.pre_base;
            pa.offtab = a.offtab;
            pa.offtab_n = a.offtab_n;
            pa.k = a.k;
            pa.item0 = P.i0;
            pa.item1 = P.i1;
            pa.out_nbrs = J.pre_nbrs;
            pa.out_cnt = J.pre_cnt;
            const uint32_t item = P.i0 + unit * kWavesPerWG + (uint32_t)wave;
            if (item < P.i1) prefetch_item(pa, wls + wave, item, lane);
        }
        // ---- the unit is done
        unit_release();
        __syncthreads();
        if (tid == 0) {
            unit_publish(spread);
            const uint64_t c2 = wall_clock64();
            s_t[P.kind == Q_RESOLVE ? 1 : 2] += c2 - c1;
            s_nu[P.kind == Q_RESOLVE ? 0 : 1] += 1;
            const uint32_t old = __hip_atomic_fetch_add(&e.dyn[jb].finished, 1u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&e.ctl->x[home].progress, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_last = old + 1u == P.units;
            if (s_last) unit_acquire();
        }
        __syncthreads();
        if (s_last) {
            const uint64_t c3 = wall_clock64();
            // closing work of the phase: every unit's results are in the L2
            if (P.kind == Q_RESOLVE && P.fuse) {
                const BatchArgs a = ld_uniform(args + P.arg);
                for (uint32_t i = P.i0 + tid; i < P.i1; i += blockDim.x)
                    commit_one(a.W, a.tiles_x, a.maps, a.work, J.result, i, P.i0, a.first_new, a.state_w, a.mask_w,
                               a.score_w, a.gmask);
            } else if (P.kind == Q_STATS && tid == 0) {
                stats_final(J.st_part, J.st_part_n, P.units, J.st_sum + P.stage, J.st_cnt + P.stage);
            }
            unit_release();
            __syncthreads();
            if (tid == 0) {
                __hip_atomic_store(&e.dyn[jb].finished, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                unit_release();
                unit_publish(spread);
                __hip_atomic_store(&e.dyn[jb].cursor, (unsigned long long)(phi + 1u) << 32, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                if (phi + 1u == J.n_phases) atomicSub(&e.ctl->x[home].active, 1u);
                s_t[3] += wall_clock64() - c3;
            }
        }
    }
    if (threadIdx.x == 0) {
        QXcd& X = e.ctl->x[xcc_id()];
        atomicAdd(&X.t_find, s_t[0]);
        atomicAdd(&X.t_resolve, s_t[1]);
        atomicAdd(&X.t_other, s_t[2]);
        atomicAdd(&X.t_close, s_t[3]);
        atomicAdd(&X.units_resolve, s_nu[0]);
        atomicAdd(&X.units_other, s_nu[1]);
    }
}

}  // namespace

namespace {
// events the next resolve launch records around its dispatch (profiling)
hipEvent_t g_ev_start = nullptr, g_ev_stop = nullptr;
}  // namespace

void set_launch_events(hipEvent_t start, hipEvent_t stop) {
    g_ev_start = start;
    g_ev_stop = stop;
}

hipError_t configure_kernels() {
    for (const void* f : {(const void*)k_engine<true>, (const void*)k_engine<false>}) {
        const hipError_t e0 = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  (int)resolve_lds_bytes(f == (const void*)k_engine<true>));
        if (e0 != hipSuccess) return e0;
    }
    hipError_t e = hipFuncSetAttribute((const void*)k_resolve<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)resolve_lds_bytes(true));
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_resolve_persist<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)resolve_lds_bytes(true));
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_resolve_persist<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)resolve_lds_bytes(false));
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_resolve<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)resolve_lds_bytes(false));
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_resolve_small<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)resolve_small_lds_bytes(true));
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_resolve_coop<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)resolve_small_lds_bytes(true));
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_resolve_coop<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)resolve_small_lds_bytes(false));
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_resolve_flow<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)resolve_lds_bytes(true));
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_resolve_flow<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)resolve_lds_bytes(false));
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)k_resolve_small<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)resolve_small_lds_bytes(false));
}

hipError_t launch_knn_prefetch(const PrefetchArgs& a, hipStream_t s) {
    const uint32_t n = a.item1 - a.item0;
    if (n == 0) return hipSuccess;
    if (a.k == 0 || a.k > (uint32_t)kMaxK) return hipErrorInvalidValue;
    if (!a.n_at && (!a.bounds || !a.bnres || a.nb == 0)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_knn_prefetch, dim3((n + kWavesPerWG - 1) / kWavesPerWG), dim3(kWave * kWavesPerWG),
                       kWavesPerWG * sizeof(WaveLds), s, a);
    return hipGetLastError();
}

hipError_t launch_unpack(const UnpackArgs& a, hipStream_t s) {
    const uint32_t n = a.n > a.n_res ? a.n : a.n_res;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpack, dim3((n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_id_maps(const uint4* state, uint32_t n, uint32_t* patch_rgba, uint32_t* map_rgba, hipStream_t s) {
    if (n == 0 || (!patch_rgba && !map_rgba)) return hipSuccess;
    hipLaunchKernelGGL(k_id_maps, dim3((n + 255) / 256), dim3(256), 0, s, state, n, patch_rgba, map_rgba);
    return hipGetLastError();
}

hipError_t launch_uncertainty(const uint32_t* rlist, const float* score, uint32_t n_res, uint32_t* out_unc,
                              hipStream_t s) {
    if (n_res == 0) return hipSuccess;
    hipLaunchKernelGGL(k_uncertainty, dim3((n_res + 255) / 256), dim3(256), 0, s, rlist, score, n_res, out_unc);
    return hipGetLastError();
}

hipError_t launch_xform_apply(const uint32_t* coord, uint32_t n, const XformSource* srcs, uint32_t n_src,
                              uint32_t* out, uint32_t* bad, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_xform_apply, dim3((n + 255) / 256), dim3(256), 0, s, coord, n, srcs, n_src, out, bad);
    return hipGetLastError();
}

hipError_t launch_rank(const uint32_t* rlist, uint32_t lo, uint32_t hi, uint32_t* rank, hipStream_t s) {
    if (hi <= lo) return hipSuccess;
    hipLaunchKernelGGL(k_rank, dim3((hi - lo + 255) / 256), dim3(256), 0, s, rlist, lo, hi, rank);
    return hipGetLastError();
}

hipError_t launch_mask_or(const uint32_t* rlist, uint32_t lo, uint32_t hi, uint32_t W, uint32_t tiles_x,
                          uint64_t* mask, uint32_t* tile_min, hipStream_t s) {
    if (hi <= lo) return hipSuccess;
    hipLaunchKernelGGL(k_mask_or, dim3((hi - lo + 255) / 256), dim3(256), 0, s, rlist, lo, hi, W, tiles_x, mask,
                       tile_min);
    return hipGetLastError();
}

hipError_t launch_mask_tiles(const uint8_t* mask, uint32_t W, uint32_t H, uint32_t tiles_x, uint32_t tiles_y,
                             uint64_t* tiles, hipStream_t s) {
    const uint32_t n = tiles_x * tiles_y;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mask_tiles, dim3((n + 255) / 256), dim3(256), 0, s, mask, W, H, tiles_x, tiles_y, tiles);
    return hipGetLastError();
}

hipError_t launch_knn_unpack(const uint32_t* q_flat, const int32_t* packed, const uint32_t* counts, uint32_t n_q,
                             uint32_t k, uint32_t W, int32_t* pairs, hipStream_t s) {
    const uint32_t n = n_q * k;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_knn_unpack, dim3((n + 255) / 256), dim3(256), 0, s, q_flat, packed, counts, n_q, k, W,
                       pairs);
    return hipGetLastError();
}

hipError_t launch_stats_chunks(const float* item_score, uint32_t n, double* part, unsigned long long* part_n,
                               double* sum, unsigned long long* count, hipStream_t s) {
    const uint32_t parts = (n + kStatsChunk - 1) / kStatsChunk;
    if (parts > 0) hipLaunchKernelGGL(k_stats_part, dim3(parts), dim3(256), 0, s, item_score, n, part, part_n);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(64), 0, s, part, part_n, parts, sum, count);
    return hipGetLastError();
}

uint32_t engine_max_wgs(bool guided) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    const void* f = guided ? (const void*)k_engine<true> : (const void*)k_engine<false>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kWave * kWavesPerWG, resolve_lds_bytes(guided)) !=
        hipSuccess)
        return 0;
    return (uint32_t)(per_cu * cus);
}

hipError_t launch_engine(const QEngineArgs& e, const BatchArgs* args, uint32_t wgs, bool guided, hipStream_t s) {
    if (wgs == 0 || !e.jobs || !e.dyn || !e.ctl || !args || e.idle_ticks == 0) return hipErrorInvalidValue;
    const size_t lds = resolve_lds_bytes(guided);
    if (guided) {
        hipLaunchKernelGGL(k_engine<true>, dim3(wgs), dim3(kWave * kWavesPerWG), lds, s, e, args);
    } else {
        hipLaunchKernelGGL(k_engine<false>, dim3(wgs), dim3(kWave * kWavesPerWG), lds, s, e, args);
    }
    return hipGetLastError();
}

hipError_t launch_exp_sweep(uint64_t i0, uint64_t n, float* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_exp_sweep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, i0, n, out);
    return hipGetLastError();
}

hipError_t launch_xcc_probe(uint32_t* out, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_xcc_probe, dim3(n), dim3(64), 0, s, out, n);
    return hipGetLastError();
}

hipError_t launch_stats(const StatsArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    uint32_t blocks = (a.n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_stats, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

size_t resolve_lds_bytes(bool guided) {
    return lut_bytes(guided) + kWavesPerWG * sizeof(WaveLds);
}

size_t resolve_small_lds_bytes(bool guided) {
    return lut_bytes(guided) + kSmallWaves * sizeof(WaveLds);
}

uint32_t small_batch_max() { return kSmallWaves; }
uint32_t items_per_wg(uint32_t team) { return kWavesPerWG / (team > 1 ? team : 1); }

// Waves per item for a batch of n items: while the batch leaves most of the
// device's `waves` resident wave slots idle, each item gets a team of up to
// team_max waves (at most a workgroup).
uint32_t team_for(uint64_t n, uint32_t team_max, uint32_t waves) {
    uint32_t T = team_max < (uint32_t)kWavesPerWG ? team_max : (uint32_t)kWavesPerWG;
    while (T > 1 && (uint64_t)T * n > (uint64_t)waves) T >>= 1;
    return T > 0 ? T : 1;
}

uint32_t coop_max_wgs(bool guided) {
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    const void* f = guided ? (const void*)k_resolve_coop<true> : (const void*)k_resolve_coop<false>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kWave * kSmallWaves,
                                                     resolve_small_lds_bytes(guided)) != hipSuccess)
        return 0;
    return (uint32_t)(per_cu * cus);
}

hipError_t launch_resolve_coop(const BatchArgs& a, const uint32_t* bounds, const uint32_t* nres, uint32_t n_batches,
                               uint32_t wgs, uint32_t team_max, uint32_t* overflow, bool guided, hipStream_t s) {
    if (n_batches == 0) return hipSuccess;
    if (wgs == 0 || wgs > coop_max_wgs(guided)) return hipErrorCooperativeLaunchTooLarge;
    BatchArgs args = a;
    void* params[] = {&args, &bounds, &nres, &n_batches, &team_max, &overflow};
    const void* f = guided ? (const void*)k_resolve_coop<true> : (const void*)k_resolve_coop<false>;
    hipError_t e;
    if (g_ev_start && (e = hipEventRecord(g_ev_start, s)) != hipSuccess) return e;
    if ((e = hipLaunchCooperativeKernel(f, dim3(wgs), dim3(kWave * kSmallWaves), params,
                                        (unsigned)resolve_small_lds_bytes(guided), s)) != hipSuccess)
        return e;
    if (g_ev_stop && (e = hipEventRecord(g_ev_stop, s)) != hipSuccess) return e;
    return hipSuccess;
}

hipError_t launch_resolve_small(const BatchArgs& a, const uint32_t* bounds, const uint32_t* nres,
                                uint32_t n_batches, bool guided, hipStream_t s) {
    if (n_batches == 0) return hipSuccess;
    const uint32_t lds = (uint32_t)resolve_small_lds_bytes(guided);
    if (guided) {
        hipExtLaunchKernelGGL(k_resolve_small<true>, dim3(1), dim3(kWave * kSmallWaves), lds, s, g_ev_start, g_ev_stop,
                              0u, a, bounds, nres, n_batches);
    } else {
        hipExtLaunchKernelGGL(k_resolve_small<false>, dim3(1), dim3(kWave * kSmallWaves), lds, s, g_ev_start,
                              g_ev_stop, 0u, a, bounds, nres, n_batches);
    }
    return hipGetLastError();
}

size_t flow_lds_bytes(uint32_t team, bool guided) {
    return lut_bytes(guided) + team * sizeof(WaveLds);
}

uint32_t persist_max_wgs(bool guided) {
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    const void* f = guided ? (const void*)k_resolve_persist<true> : (const void*)k_resolve_persist<false>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kWave * kWavesPerWG, resolve_lds_bytes(guided)) !=
        hipSuccess)
        return 0;
    return (uint32_t)(per_cu * cus);
}

hipError_t launch_resolve_persist(const BatchArgs& a, uint32_t wgs, bool guided, hipStream_t s) {
    if (a.item1 <= a.item0) return hipSuccess;
    if (wgs == 0 || a.team > 1 || a.fused_commit || a.state2) return hipErrorInvalidValue;
    const uint32_t lds = (uint32_t)resolve_lds_bytes(guided);
    if (guided) {
        hipExtLaunchKernelGGL(k_resolve_persist<true>, dim3(wgs), dim3(kWave * kWavesPerWG), lds, s, g_ev_start,
                              g_ev_stop, 0u, a);
    } else {
        hipExtLaunchKernelGGL(k_resolve_persist<false>, dim3(wgs), dim3(kWave * kWavesPerWG), lds, s, g_ev_start,
                              g_ev_stop, 0u, a);
    }
    return hipGetLastError();
}

uint32_t flow_max_wgs(uint32_t team, bool guided) {
    if (team == 0 || team > (uint32_t)kWavesPerWG) return 0;
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    const void* f = guided ? (const void*)k_resolve_flow<true> : (const void*)k_resolve_flow<false>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kWave * team, flow_lds_bytes(team, guided)) !=
        hipSuccess)
        return 0;
    return (uint32_t)(per_cu * cus);
}

hipError_t launch_resolve_flow(const BatchArgs& a, const FlowArgs& f, uint32_t wgs, uint32_t team, bool guided,
                               hipStream_t s) {
    if (f.n == 0) return hipSuccess;
    if (team == 0 || team > (uint32_t)kWavesPerWG || (team & (team - 1)) != 0) return hipErrorInvalidValue;
    if (wgs == 0 || !f.ticket || !f.dep || !f.flag || !f.overflow || f.epoch == 0 || f.spin_limit == 0)
        return hipErrorInvalidValue;
    // every item needs its neighbour list; test hooks are not supported
    if (!a.pre_nbrs || a.item0 < a.pre_base || a.item0 + f.n > a.pre_end || a.ext_nbrs || a.ext_cands ||
        a.dbg_nbrs || a.dbg_best || a.knn_only || a.ext_full_cost || a.gmask)
        return hipErrorInvalidValue;
    const uint32_t lds = (uint32_t)flow_lds_bytes(team, guided);
    if (guided) {
        hipExtLaunchKernelGGL(k_resolve_flow<true>, dim3(wgs), dim3(kWave * team), lds, s, g_ev_start, g_ev_stop, 0u,
                              a, f);
    } else {
        hipExtLaunchKernelGGL(k_resolve_flow<false>, dim3(wgs), dim3(kWave * team), lds, s, g_ev_start, g_ev_stop,
                              0u, a, f);
    }
    return hipGetLastError();
}

hipError_t launch_resolve(const BatchArgs& a, bool guided, hipStream_t s) {
    uint32_t n = a.item1 - a.item0;
    if (n == 0) return hipSuccess;
    if (a.team > kWavesPerWG || (a.team & (a.team - 1)) != 0) return hipErrorInvalidValue;
    const uint32_t ipw = items_per_wg(a.team);
    dim3 grid((n + ipw - 1) / ipw), block(kWave * kWavesPerWG);
    const uint32_t lds = (uint32_t)resolve_lds_bytes(guided);
    if (guided) {
        hipExtLaunchKernelGGL(k_resolve<true>, grid, block, lds, s, g_ev_start, g_ev_stop, 0u, a);
    } else {
        hipExtLaunchKernelGGL(k_resolve<false>, grid, block, lds, s, g_ev_start, g_ev_stop, 0u, a);
    }
    return hipGetLastError();
}

hipError_t launch_rand(const RandArgs& a, hipStream_t s) {
    uint32_t n = a.item1 - a.item0;
    if (n == 0 || a.m == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rand, dim3((n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_commit(const CommitArgs& a, hipStream_t s) {
    uint32_t n = a.item1 - a.item0;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_commit, dim3((n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_defer_flush(const DeferFlushArgs& a, hipStream_t s) {
    const uint32_t n0 = a.prev1 - a.prev0, n1 = a.pp1 - a.pp0, n = n0 > n1 ? n0 : n1;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_defer_flush, dim3((n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_state_snapshot(const uint32_t* rlist, uint32_t n, const uint4* state, uint4* old, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_state_snapshot, dim3((n + 255) / 256), dim3(256), 0, s, rlist, n, state, old);
    return hipGetLastError();
}

hipError_t launch_recolor(const RecolorArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_recolor, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_resample_v(const uint32_t* src, uint32_t w, uint32_t h, float4* dst, uint32_t nh,
                             const uint32_t* left, const uint32_t* cnt, const float* ws, uint32_t stride,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_resample_v, dim3((w + 255) / 256, nh), dim3(256), 0, s, src, w, h, dst, nh, left, cnt, ws,
                       stride);
    return hipGetLastError();
}

hipError_t launch_resample_h(const float4* src, uint32_t w, uint32_t h, uint32_t* dst, uint32_t nw,
                             const uint32_t* left, const uint32_t* cnt, const float* ws, uint32_t stride,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_resample_h, dim3((nw + 255) / 256, h), dim3(256), 0, s, src, w, h, dst, nw, left, cnt, ws,
                       stride);
    return hipGetLastError();
}

}  // namespace tsg
