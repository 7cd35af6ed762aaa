/*
 * refcpu — CPU restatement of texture-synthesis's per-pixel patch search.
 *
 * TEST INFRASTRUCTURE: the correctness oracle and the timed CPU baseline.
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg. The product library never links or calls it.
 *
 * Each function restates the reference function cited beside it
 * (EmbarkStudios/texture-synthesis, lib/src/...). Third-party arithmetic the
 * reference calls is restated from the published algorithms of the pinned
 * crates (Cargo.lock): rand_pcg 0.3.1, rand_core 0.6.3, rand 0.8.5,
 * image 0.23.12. rstar 0.7.1's order among equidistant points is not
 * published; this oracle declares the canonical order (d^2, dy, dx).
 *
 * One extension beyond the reference: work items of a stage are committed
 * in Jacobi batches (all items of a batch read the state as it was at the
 * batch start). A batch size of 1 is exactly the reference's
 * max_thread_count(1) behaviour.
 *
 * Build with -ffp-contract=off: the reference's f32 sums are separate
 * multiplies and adds.
 */
#include "refcpu.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

/* ------------------------------------------------------------------ */
/* rand_pcg 0.3.1 Lcg64Xsh32 (PCG-XSH-RR 64/32)                        */
/* ------------------------------------------------------------------ */
constexpr uint64_t kPcgMul = 6364136223846793005ULL;
constexpr uint64_t kSeedInc = 11634580027462260723ULL;

inline uint32_t rotr32(uint32_t x, uint32_t r) { return (x >> r) | (x << ((32u - r) & 31u)); }

struct Pcg32 {
    uint64_t state, inc;
    void step() { state = state * kPcgMul + inc; }
    /* next_u32: output from the pre-advance state */
    uint32_t next_u32() {
        uint64_t s = state;
        step();
        uint32_t rot = (uint32_t)(s >> 59);
        uint32_t xsh = (uint32_t)(((s >> 18) ^ s) >> 27);
        return rotr32(xsh, rot);
    }
    /* rand_core::impls::next_u64_via_u32: low word first */
    uint64_t next_u64() {
        uint64_t lo = next_u32();
        uint64_t hi = next_u32();
        return (hi << 32) | lo;
    }
    static Pcg32 from_state_incr(uint64_t st, uint64_t inc) {
        Pcg32 p{st, inc};
        p.state = p.state + p.inc;
        p.step();
        return p;
    }
    static Pcg32 make(uint64_t st, uint64_t stream) { return from_state_incr(st, (stream << 1) | 1u); }
    static Pcg32 from_seed(const uint8_t b[16]) {
        uint64_t s0 = 0, s1 = 0;
        for (int i = 7; i >= 0; --i) s0 = (s0 << 8) | b[i];
        for (int i = 15; i >= 8; --i) s1 = (s1 << 8) | b[i];
        return from_state_incr(s0, s1 | 1u);
    }
    /* rand_core 0.6.3 SeedableRng::seed_from_u64 default: four PCG32
     * outputs of a private LCG, advance-then-output, little-endian bytes. */
    static Pcg32 seed_from_u64(uint64_t st) {
        uint8_t seed[16];
        for (int c = 0; c < 4; ++c) {
            st = st * kPcgMul + kSeedInc;
            uint64_t s = st;
            uint32_t x = rotr32((uint32_t)(((s >> 18) ^ s) >> 27), (uint32_t)(s >> 59));
            for (int b = 0; b < 4; ++b) seed[c * 4 + b] = (uint8_t)(x >> (8 * b));
        }
        return from_seed(seed);
    }
};

/* rand 0.8.5 UniformInt::<u32>::sample_single(lo, hi) ->
 * sample_single_inclusive(lo, hi-1): Lemire widening multiply with the
 * "(range << lz) - 1" zone. */
inline uint32_t range_u32(Pcg32& g, uint32_t lo, uint32_t hi) {
    uint32_t range = hi - lo;
    if (range == 0) return g.next_u32();
    uint32_t zone = (range << __builtin_clz(range)) - 1u;
    for (;;) {
        uint32_t v = g.next_u32();
        uint64_t m = (uint64_t)v * range;
        if ((uint32_t)m <= zone) return lo + (uint32_t)(m >> 32);
    }
}
/* rand 0.8.5 UniformInt::<u8 / u16>::sample_single: u32 draws; for types of
 * at most 16 bits the zone is u32::MAX - (u32::MAX - range + 1) % range. */
inline uint32_t range_u8(Pcg32& g, uint32_t lo, uint32_t hi) {
    const uint32_t range = hi - lo;
    const uint32_t unsigned_max = 0xffffffffu;
    const uint32_t zone = unsigned_max - (unsigned_max - range + 1u) % range;
    for (;;) {
        uint32_t v = g.next_u32();
        uint64_t m = (uint64_t)v * range;
        if ((uint32_t)m <= zone) return lo + (uint32_t)(m >> 32);
    }
}
/* same for u64 / usize (x86-64): draws next_u64 */
inline uint64_t range_u64(Pcg32& g, uint64_t lo, uint64_t hi) {
    uint64_t range = hi - lo;
    if (range == 0) return g.next_u64();
    uint64_t zone = (range << __builtin_clzll(range)) - 1u;
    for (;;) {
        uint64_t v = g.next_u64();
        unsigned __int128 m = (unsigned __int128)v * range;
        if ((uint64_t)m <= zone) return lo + (uint64_t)(m >> 64);
    }
}

/* ------------------------------------------------------------------ */
/* image 0.23.12 imageops::resize (sample.rs), restated                */
/* ------------------------------------------------------------------ */
float k_triangle(float x) {
    float a = std::fabs(x);
    return a < 1.0f ? 1.0f - a : 0.0f;
}
float k_catmullrom(float x) { /* bc_cubic_spline(x, 0.0, 0.5) */
    const float b = 0.0f, c = 0.5f;
    float a = std::fabs(x);
    float k;
    if (a < 1.0f) {
        k = (12.0f - 9.0f * b - 6.0f * c) * (a * a * a) + (-18.0f + 12.0f * b + 6.0f * c) * (a * a) + (6.0f - 2.0f * b);
    } else if (a < 2.0f) {
        k = (-b - 6.0f * c) * (a * a * a) + (6.0f * b + 30.0f * c) * (a * a) + (-12.0f * b - 48.0f * c) * a + (8.0f * b + 24.0f * c);
    } else {
        k = 0.0f;
    }
    return k / 6.0f;
}
float k_gaussian(float x) { /* gaussian(x, 0.5) */
    const float r = 0.5f;
    float c = 1.0f / (std::sqrt(2.0f * 3.14159265358979323846f) * r);
    return c * std::exp(-(x * x) / (2.0f * (r * r)));
}

struct Filter {
    float (*kernel)(float);
    float support;
};
/* imageops::blur(image, 2.0): gaussian(x, sigma) with support 2 sigma */
float k_gaussian_blur2(float x) {
    const float r = 2.0f;
    float c = 1.0f / (std::sqrt(2.0f * 3.14159265358979323846f) * r);
    return c * std::exp(-(x * x) / (2.0f * (r * r)));
}
Filter filter_of(int f) {
    if (f == 0) return {k_triangle, 1.0f};
    if (f == 1) return {k_catmullrom, 2.0f};
    if (f == 3) return {k_gaussian_blur2, 4.0f};
    return {k_gaussian, 3.0f};
}

/* per output index: [left, right) and normalised weights */
void sample_weights(uint32_t in_len, uint32_t out_len, const Filter& flt, std::vector<uint32_t>& left,
                    std::vector<uint32_t>& cnt, std::vector<float>& ws, uint32_t& stride) {
    float ratio = (float)in_len / (float)out_len;
    float sratio = ratio < 1.0f ? 1.0f : ratio;
    float src_support = flt.support * sratio;
    left.resize(out_len);
    cnt.resize(out_len);
    std::vector<std::vector<float>> tmp(out_len);
    stride = 0;
    for (uint32_t o = 0; o < out_len; ++o) {
        float input = ((float)o + 0.5f) * ratio;
        int64_t l = (int64_t)std::floor(input - src_support);
        l = std::min<int64_t>(std::max<int64_t>(l, 0), (int64_t)in_len - 1);
        int64_t r = (int64_t)std::ceil(input + src_support);
        r = std::min<int64_t>(std::max<int64_t>(r, l + 1), (int64_t)in_len);
        input = input - 0.5f;
        float sum = 0.0f;
        for (int64_t i = l; i < r; ++i) {
            float w = flt.kernel(((float)i - input) / sratio);
            tmp[o].push_back(w);
            sum += w;
        }
        for (float& w : tmp[o]) w /= sum;
        left[o] = (uint32_t)l;
        cnt[o] = (uint32_t)(r - l);
        stride = std::max(stride, cnt[o]);
    }
    ws.assign((size_t)out_len * stride, 0.0f);
    for (uint32_t o = 0; o < out_len; ++o)
        for (uint32_t i = 0; i < cnt[o]; ++i) ws[(size_t)o * stride + i] = tmp[o][i];
}

/* vertical_sample then horizontal_sample; f32 intermediate; rounded u8 out */
void resize_rgba(const uint8_t* src, uint32_t w, uint32_t h, uint32_t nw, uint32_t nh, int filter, uint8_t* dst) {
    Filter flt = filter_of(filter);
    std::vector<uint32_t> vl, vc, hl, hc;
    std::vector<float> vw, hw;
    uint32_t vs, hs;
    sample_weights(h, nh, flt, vl, vc, vw, vs);
    sample_weights(w, nw, flt, hl, hc, hw, hs);
    std::vector<float> tmp((size_t)w * nh * 4);
    for (uint32_t oy = 0; oy < nh; ++oy) {
        for (uint32_t x = 0; x < w; ++x) {
            float t[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (uint32_t i = 0; i < vc[oy]; ++i) {
                float wt = vw[(size_t)oy * vs + i];
                const uint8_t* p = src + ((size_t)(vl[oy] + i) * w + x) * 4;
                for (int c = 0; c < 4; ++c) t[c] += (float)p[c] * wt;
            }
            for (int c = 0; c < 4; ++c) tmp[((size_t)oy * w + x) * 4 + c] = t[c];
        }
    }
    for (uint32_t ox = 0; ox < nw; ++ox) {
        for (uint32_t y = 0; y < nh; ++y) {
            float t[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (uint32_t i = 0; i < hc[ox]; ++i) {
                float wt = hw[(size_t)ox * hs + i];
                const float* p = &tmp[((size_t)y * w + hl[ox] + i) * 4];
                for (int c = 0; c < 4; ++c) t[c] += p[c] * wt;
            }
            for (int c = 0; c < 4; ++c) {
                float v = t[c] < 0.0f ? 0.0f : (t[c] > 255.0f ? 255.0f : t[c]);
                dst[((size_t)y * nw + ox) * 4 + c] = (uint8_t)std::round(v);
            }
        }
    }
}

struct Img {
    uint32_t w = 0, h = 0;
    std::vector<uint8_t> px;
    const uint8_t* at(uint32_t x, uint32_t y) const { return &px[((size_t)y * w + x) * 4]; }
    bool in_bounds(int32_t x, int32_t y) const { return x >= 0 && y >= 0 && x < (int32_t)w && y < (int32_t)h; }
};

/* img_pyramid.rs:20-37 build_gaussian: levels-1 blurred levels, coarsest
 * first, each resize(resize(img, w/2^i, h/2^i, Gaussian), w, h, Gaussian),
 * then the input itself. */
bool build_pyramid(const Img& in, uint32_t levels, std::vector<Img>& out) {
    out.clear();
    for (int i = (int)levels - 1; i >= 1; --i) {
        uint32_t pw = in.w >> i, ph = in.h >> i;
        if (pw == 0 || ph == 0) return false;
        std::vector<uint8_t> small((size_t)pw * ph * 4);
        resize_rgba(in.px.data(), in.w, in.h, pw, ph, 2, small.data());
        Img lv;
        lv.w = in.w;
        lv.h = in.h;
        lv.px.resize((size_t)in.w * in.h * 4);
        resize_rgba(small.data(), pw, ph, in.w, in.h, 2, lv.px.data());
        out.push_back(std::move(lv));
    }
    out.push_back(in);
    return true;
}

/* ------------------------------------------------------------------ */
/* cost tables (ms.rs:1109-1120, 1290-1311, 739-742, 853-858)          */
/* ------------------------------------------------------------------ */
inline float metric_cauchy(uint8_t a, uint8_t b, float sig2) {
    float x2 = ((float)a - (float)b) / 255.0f;
    x2 = x2 * x2;
    return std::log1p(x2 / sig2);
}
inline float metric_l2(uint8_t a, uint8_t b) {
    float x = ((float)a - (float)b) / 255.0f;
    return x * x;
}

/* ------------------------------------------------------------------ */
/* canonical exact kNN                                                 */
/* ------------------------------------------------------------------ */
struct Off {
    int32_t dx, dy;
};
constexpr int kTableRadius = 64;
const std::vector<Off>& offset_table() {
    static std::vector<Off> t = [] {
        std::vector<Off> v;
        for (int dy = -kTableRadius; dy <= kTableRadius; ++dy)
            for (int dx = -kTableRadius; dx <= kTableRadius; ++dx)
                if (dx * dx + dy * dy <= kTableRadius * kTableRadius) v.push_back({dx, dy});
        std::sort(v.begin(), v.end(), [](const Off& a, const Off& b) {
            int da = a.dx * a.dx + a.dy * a.dy, db = b.dx * b.dx + b.dy * b.dy;
            if (da != db) return da < db;
            if (a.dy != b.dy) return a.dy < b.dy;
            return a.dx < b.dx;
        });
        return v;
    }();
    return t;
}
inline bool key_less(int64_t d2a, int32_t dya, int32_t dxa, int64_t d2b, int32_t dyb, int32_t dxb) {
    if (d2a != d2b) return d2a < d2b;
    if (dya != dyb) return dya < dyb;
    return dxa < dxb;
}

/* Resolved-point field with the tiling-mode ghosts of flush_resolved
 * (ms.rs:296-331): a pixel flushed in tiling mode within 5% of an edge is
 * also present shifted by +-W (x) and by +-H (y); no diagonal copies. */
struct PointField {
    const uint8_t* mask;   /* resolved */
    const uint8_t* ghost;  /* flushed with ghosts (NULL when not tiling) */
    int32_t W, H, xl, xr, yb, yt;
    bool exists(int32_t X, int32_t Y) const {
        bool xin = X >= 0 && X < W, yin = Y >= 0 && Y < H;
        if (xin && yin) return mask[(size_t)Y * W + X] != 0;
        if (!ghost || (!xin && !yin)) return false;
        if (!xin) {
            if (X >= W) {
                int32_t x0 = X - W;
                return x0 < W && x0 < xl && ghost[(size_t)Y * W + x0];
            }
            int32_t x0 = X + W;
            return x0 >= 0 && x0 > xr && ghost[(size_t)Y * W + x0];
        }
        if (Y >= H) {
            int32_t y0 = Y - H;
            return y0 < H && y0 < yb && ghost[(size_t)y0 * W + X];
        }
        int32_t y0 = Y + H;
        return y0 >= 0 && y0 > yt && ghost[(size_t)y0 * W + X];
    }
};

struct KnnEntry {
    int64_t d2;
    int32_t dy, dx;
};

/* k nearest existing points to (x,y) in canonical order. n_points is the
 * number of resolved pixels (used only to choose the search route; both
 * routes return the same list). */
void knn_query(const PointField& f, const std::vector<uint32_t>& resolved_list, int32_t x, int32_t y, uint32_t k,
               std::vector<int32_t>& nx, std::vector<int32_t>& ny) {
    nx.clear();
    ny.clear();
    if (k == 0) return;
    size_t n_points = resolved_list.size();
    if (n_points >= 256) {
        for (const Off& o : offset_table()) {
            if (f.exists(x + o.dx, y + o.dy)) {
                nx.push_back(x + o.dx);
                ny.push_back(y + o.dy);
                if (nx.size() == k) return;
            }
        }
        nx.clear();
        ny.clear();
    }
    /* exhaustive route: every resolved pixel and its ghosts */
    std::vector<KnnEntry> all;
    all.reserve(n_points * (f.ghost ? 2 : 1));
    auto add = [&](int32_t X, int32_t Y) {
        int64_t dx = X - x, dy = Y - y;
        all.push_back({dx * dx + dy * dy, (int32_t)dy, (int32_t)dx});
    };
    for (uint32_t flat : resolved_list) {
        int32_t px = (int32_t)(flat % (uint32_t)f.W), py = (int32_t)(flat / (uint32_t)f.W);
        if (!f.mask[flat]) continue;
        add(px, py);
        if (f.ghost && f.ghost[flat]) {
            if (px < f.xl) add(px + f.W, py);
            else if (px > f.xr) add(px - f.W, py);
            if (py < f.yb) add(px, py + f.H);
            else if (py > f.yt) add(px, py - f.H);
        }
    }
    size_t take = std::min<size_t>(k, all.size());
    std::partial_sort(all.begin(), all.begin() + take, all.end(), [](const KnnEntry& a, const KnnEntry& b) {
        return key_less(a.d2, a.dy, a.dx, b.d2, b.dy, b.dx);
    });
    for (size_t i = 0; i < take; ++i) {
        nx.push_back(x + all[i].dx);
        ny.push_back(y + all[i].dy);
    }
}

template <class F>
void parallel_for(size_t n, unsigned threads, F&& f) {
    if (threads <= 1 || n < 2 * (size_t)threads) {
        for (size_t i = 0; i < n; ++i) f(i, 0u);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> th;
    for (unsigned t = 0; t < threads; ++t)
        th.emplace_back([&, t] {
            for (;;) {
                size_t i = next.fetch_add(1, std::memory_order_relaxed);
                if (i >= n) break;
                f(i, t);
            }
        });
    for (auto& x : th) x.join();
}

inline int32_t modulo(int32_t a, int32_t b) {
    int32_t r = a % b;
    return r < 0 ? r + b : r;
}

const uint8_t kOutside[4] = {0, 0, 0, 255};

}  // namespace

/* ------------------------------------------------------------------ */
/* the synthesis job                                                   */
/* ------------------------------------------------------------------ */
struct Cand {
    int32_t x, y;
    uint32_t map, kind, patch, pmap;
    std::vector<int32_t> kx, ky; /* k_neighs of CandidateStruct */
};

struct Scratch {
    std::vector<int32_t> nx, ny;
    std::vector<double> dist;
    std::vector<float> w;
    std::vector<Cand> cands;
    size_t n_cands = 0;
    std::vector<uint8_t> pat, gpat;
    bool error = false;
};

struct Result {
    uint32_t flat;
    bool is_new, fallback;
    uint32_t sx, sy, smap, patch, pmap;
    float score;
};

struct orc_job {
    orc_params prm{};
    uint32_t W = 0, H = 0;
    size_t N0 = 0;
    uint32_t levels = 1;
    std::vector<std::vector<Img>> ex;  /* unfiltered examples, [example][level] */
    std::vector<int> kind;
    std::vector<Img> smask;
    std::vector<uint32_t> filt;        /* filtered map id -> example index */
    bool guided = false;
    std::vector<std::vector<Img>> exg; /* unfiltered example guides */
    std::vector<Img> tg;               /* target guide pyramid */
    /* state (ms.rs:207-217) */
    std::vector<uint8_t> color;
    std::vector<uint32_t> cx, cy, cm, pid, pm;
    std::vector<uint32_t> unresolved;
    std::vector<uint32_t> resolved;
    std::vector<float> rscore;
    size_t locked = 0;
    std::vector<uint8_t> mask, ghost;
    /* run cursor */
    bool initialized = false, finished = false;
    int t = 0;            /* stage index; p_stage = P - t */
    bool stage_open = false;
    uint32_t level = 0;
    float alpha_s = 0.0f;
    uint64_t pss = 0, S = 0, redo = 0, cursor = 0;
    bool stage_exhausted = false;
    std::vector<float> my_full, g_full;
    std::vector<const Img*> maps;        /* filtered maps at the stage level */
    std::vector<const Img*> gmaps;       /* unfiltered example guides at level */
    const Img* tgl = nullptr;
    std::vector<const uint8_t*> smask_f; /* filtered sample masks (NULL = All) */
    std::vector<double> ssum;
    std::vector<uint64_t> scnt, sitems;
    std::string err;
    orc_sched sched{};
    unsigned threads = 1;
    std::vector<Scratch> scratch;
};

namespace {

bool set_err(orc_job* j, const char* m) {
    if (j->err.empty()) j->err = m;
    return false;
}

PointField field_of(const orc_job* j) {
    PointField f;
    f.mask = j->mask.data();
    f.ghost = j->prm.tiling_mode ? j->ghost.data() : nullptr;
    f.W = (int32_t)j->W;
    f.H = (int32_t)j->H;
    f.xl = (int32_t)((float)j->W * 0.05f);
    f.xr = (int32_t)j->W - f.xl;
    f.yb = (int32_t)((float)j->H * 0.05f);
    f.yt = (int32_t)j->H - f.yb;
    return f;
}

/* check_coord_validity (ms.rs:1533-1549) */
inline bool coord_valid(const orc_job* j, int32_t x, int32_t y, uint32_t map, bool& err) {
    if (map >= j->maps.size()) {
        err = true;
        return false;
    }
    const Img* m = j->maps[map];
    if (!m->in_bounds(x, y)) return false;
    const uint8_t* sm = j->smask_f[map];
    if (!sm) return true;
    return sm[((size_t)y * m->w + (uint32_t)x) * 4] != 0;
}

/* update (ms.rs:334-377) minus the list push, which commit() does */
void write_pixel(orc_job* j, uint32_t flat, uint32_t sx, uint32_t sy, uint32_t smap, uint32_t patch, uint32_t pmap,
                 const std::vector<const Img*>& maps) {
    j->cx[flat] = sx;
    j->cy[flat] = sy;
    j->cm[flat] = smap;
    j->pid[flat] = patch;
    j->pm[flat] = pmap;
    const uint8_t* c = maps[smap]->at(sx, sy);
    memcpy(&j->color[(size_t)flat * 4], c, 4);
}

/* flush_resolved (ms.rs:296-331) for one pixel */
void flush(orc_job* j, uint32_t flat, float score, bool tiling_ghosts) {
    j->mask[flat] = 1;
    if (tiling_ghosts) j->ghost[flat] = 1;
    j->resolved.push_back(flat);
    j->rscore.push_back(score);
}

/* resolve_at_random (ms.rs:447-475) source selection */
void random_source(uint64_t seed, const std::vector<const Img*>& maps, uint32_t& rmap, uint32_t& rx, uint32_t& ry) {
    Pcg32 a = Pcg32::seed_from_u64(seed);
    rmap = (uint32_t)range_u64(a, 0, maps.size());
    Pcg32 b = Pcg32::seed_from_u64(seed);
    rx = range_u32(b, 0, maps[rmap]->w);
    Pcg32 c = Pcg32::seed_from_u64(seed);
    ry = range_u32(c, 0, maps[rmap]->h);
}

/* pick_random_unresolved (ms.rs:380-389) */
bool pick_unresolved(orc_job* j, uint64_t seed, uint32_t& out) {
    if (j->unresolved.empty()) return false;
    Pcg32 g = Pcg32::seed_from_u64(seed);
    uint64_t idx = range_u64(g, 0, j->unresolved.size());
    out = j->unresolved[idx];
    j->unresolved[idx] = j->unresolved.back();
    j->unresolved.pop_back();
    return true;
}

/* get_distances_to_k_neighs (ms.rs:405-425) + the exp of find_best_match
 * (ms.rs:1198-1203) */
void weights_of(const orc_job* j, Scratch& s, int32_t x, int32_t y) {
    size_t kf = s.nx.size();
    double dimx = (double)j->W, dimy = (double)j->H;
    double x2 = (double)x / dimx, y2 = (double)y / dimy;
    s.dist.resize(kf * 4);
    for (size_t q = 0; q < kf; ++q) {
        double x1 = (double)s.nx[q] / dimx, y1 = (double)s.ny[q] / dimy;
        double d = std::fma(x1 - x2, x1 - x2, (y1 - y2) * (y1 - y2));
        s.dist[q * 4 + 0] = d;
        s.dist[q * 4 + 1] = d;
        s.dist[q * 4 + 2] = d;
        s.dist[q * 4 + 3] = d;
    }
    double sum = 0.0;
    for (double d : s.dist) sum += d;
    double avg = sum / (double)s.dist.size();
    for (double& d : s.dist) d /= avg;
    s.w.resize(kf);
    for (size_t q = 0; q < kf; ++q) s.w[q] = (float)std::exp(-1.0 * s.dist[q * 4]);
}

/* A neighbour's record as find_candidates reads it: source coordinate and
 * patch id / map (coord_map and id_map of ms.rs:207-217). */
struct NbrRec {
    uint32_t cx, cy, pid, pm;
};
struct PlainRead {
    const orc_job* j;
    NbrRec operator()(uint32_t f) const { return {j->cx[f], j->cy[f], j->pid[f], j->pm[f]}; }
};

/* find_candidates (ms.rs:478-602) */
template <class Rd>
void find_candidates(const orc_job* j, Scratch& s, int32_t px, int32_t py, uint64_t m_seed, Rd rd) {
    size_t kf = s.nx.size();
    size_t count = 0;
    int32_t W = (int32_t)j->W, H = (int32_t)j->H;
    size_t need = kf + (size_t)j->prm.random_sample_locations;
    if (s.cands.size() < need) s.cands.resize(need);
    for (size_t q = 0; q < kf; ++q) {
        int32_t shx = px - s.nx[q], shy = py - s.ny[q];
        uint32_t nflat = (uint32_t)modulo(s.ny[q], H) * (uint32_t)W + (uint32_t)modulo(s.nx[q], W);
        const NbrRec nr = rd(nflat);
        int32_t ccx = (int32_t)nr.cx + shx, ccy = (int32_t)nr.cy + shy;
        uint32_t npid = nr.pid, nmap = nr.pm;
        if (coord_valid(j, ccx, ccy, nmap, s.error)) {
            Cand& c = s.cands[count];
            c.kx.resize(kf);
            c.ky.resize(kf);
            for (size_t r = 0; r < kf; ++r) {
                c.kx[r] = ccx + (s.nx[r] - px);
                c.ky[r] = ccy + (s.ny[r] - py);
            }
            c.x = ccx;
            c.y = ccy;
            c.map = nmap;
            c.kind = 0;
            c.patch = npid;
            c.pmap = nmap;
            ++count;
        }
    }
    Pcg32 rng = Pcg32::seed_from_u64(m_seed);
    size_t nmaps = j->maps.size();
    for (uint64_t r = 0; r < j->prm.random_sample_locations; ++r) {
        uint32_t rmap = (uint32_t)range_u64(rng, 0, nmaps);
        uint32_t w = j->maps[rmap]->w, h = j->maps[rmap]->h;
        int32_t rx, ry;
        for (;;) {
            rx = (int32_t)range_u32(rng, 0, w);
            ry = (int32_t)range_u32(rng, 0, h);
            if (coord_valid(j, rx, ry, rmap, s.error)) break;
        }
        Cand& c = s.cands[count];
        c.kx.resize(kf);
        c.ky.resize(kf);
        for (size_t q = 0; q < kf; ++q) {
            c.kx[q] = rx + (px - s.nx[q]);
            c.ky[q] = ry + (py - s.ny[q]);
        }
        c.x = rx;
        c.y = ry;
        c.map = rmap;
        c.kind = 1;
        c.patch = (uint32_t)ry * w + (uint32_t)rx;
        c.pmap = rmap;
        ++count;
    }
    s.n_cands = count;
}

/* k_neighs_to_precomputed_reference_pattern (ms.rs:1151-1181) */
void pattern_of(const Img& src, const Scratch& s, bool wrap, std::vector<uint8_t>& pat) {
    size_t kf = s.nx.size();
    pat.resize(kf * 4);
    for (size_t q = 0; q < kf; ++q) {
        int32_t x = s.nx[q], y = s.ny[q];
        if (wrap) {
            x = modulo(x, (int32_t)src.w);
            y = modulo(y, (int32_t)src.h);
        }
        const uint8_t* c = src.in_bounds(x, y) ? src.at((uint32_t)x, (uint32_t)y) : kOutside;
        memcpy(&pat[q * 4], c, 4);
    }
}

/* better_match (ms.rs:1227-1288) */
bool better_match(const orc_job* j, const Scratch& s, const Cand& c, float current_best, float* out) {
    float score = 0.0f;
    size_t kf = s.nx.size();
    const Img* m = j->maps[c.map];
    const Img* g = j->guided ? j->gmaps[c.map] : nullptr;
    for (size_t q = 0; q < kf; ++q) {
        float nps = 0.0f;
        const uint8_t* p = m->in_bounds(c.kx[q], c.ky[q]) ? m->at((uint32_t)c.kx[q], (uint32_t)c.ky[q]) : kOutside;
        for (int ch = 0; ch < 4; ++ch) nps += j->my_full[(size_t)s.pat[q * 4 + ch] * 256 + p[ch]];
        if (g) {
            const uint8_t* gp = g->in_bounds(c.kx[q], c.ky[q]) ? g->at((uint32_t)c.kx[q], (uint32_t)c.ky[q]) : kOutside;
            for (int ch = 0; ch < 4; ++ch) nps += j->g_full[(size_t)s.gpat[q * 4 + ch] * 256 + gp[ch]];
        }
        score += nps * s.w[q];
        if (score >= current_best) return false;
    }
    *out = score;
    return true;
}

float full_cost(const orc_job* j, const Scratch& s, const Cand& c) {
    float score = 0.0f;
    size_t kf = s.nx.size();
    const Img* m = j->maps[c.map];
    const Img* g = j->guided ? j->gmaps[c.map] : nullptr;
    for (size_t q = 0; q < kf; ++q) {
        float nps = 0.0f;
        const uint8_t* p = m->in_bounds(c.kx[q], c.ky[q]) ? m->at((uint32_t)c.kx[q], (uint32_t)c.ky[q]) : kOutside;
        for (int ch = 0; ch < 4; ++ch) nps += j->my_full[(size_t)s.pat[q * 4 + ch] * 256 + p[ch]];
        if (g) {
            const uint8_t* gp = g->in_bounds(c.kx[q], c.ky[q]) ? g->at((uint32_t)c.kx[q], (uint32_t)c.ky[q]) : kOutside;
            for (int ch = 0; ch < 4; ++ch) nps += j->g_full[(size_t)s.gpat[q * 4 + ch] * 256 + gp[ch]];
        }
        score += nps * s.w[q];
    }
    return score;
}

/* find_best_match (ms.rs:1184-1224) */
void find_best_match(const orc_job* j, const Scratch& s, size_t& best, float& lowest) {
    best = 0;
    lowest = FLT_MAX;
    for (size_t i = 0; i < s.n_cands; ++i) {
        float c;
        if (better_match(j, s, s.cands[i], lowest, &c)) {
            lowest = c;
            best = i;
        }
    }
}

/* steps 2-4 of the worker loop (ms.rs:925-1009) against the frozen state */
void evaluate(const orc_job* j, Scratch& s, uint32_t flat, bool is_new, uint64_t loop_seed, const PointField& f,
              Result& r) {
    int32_t x = (int32_t)(flat % j->W), y = (int32_t)(flat / j->W);
    r.flat = flat;
    r.is_new = is_new;
    knn_query(f, j->resolved, x, y, j->prm.nearest_neighbors, s.nx, s.ny);
    if (s.nx.empty()) {
        r.fallback = true;
        random_source(j->pss, j->maps, r.smap, r.sx, r.sy);
        r.patch = flat;
        r.pmap = r.smap;
        r.score = 0.0f;
        return;
    }
    r.fallback = false;
    weights_of(j, s, x, y);
    find_candidates(j, s, x, y, loop_seed + 1, PlainRead{j});
    /* own pattern from the output colour map; wrap iff tiling */
    {
        size_t kf = s.nx.size();
        s.pat.resize(kf * 4);
        for (size_t q = 0; q < kf; ++q) {
            int32_t nx = s.nx[q], ny = s.ny[q];
            if (j->prm.tiling_mode) {
                nx = modulo(nx, (int32_t)j->W);
                ny = modulo(ny, (int32_t)j->H);
            }
            bool in = nx >= 0 && ny >= 0 && nx < (int32_t)j->W && ny < (int32_t)j->H;
            const uint8_t* c = in ? &j->color[((size_t)ny * j->W + nx) * 4] : kOutside;
            memcpy(&s.pat[q * 4], c, 4);
        }
    }
    if (j->guided) pattern_of(*j->tgl, s, j->prm.tiling_mode != 0, s.gpat);
    size_t best;
    float lowest;
    find_best_match(j, s, best, lowest);
    const Cand& c = s.cands[best];
    r.sx = (uint32_t)c.x;
    r.sy = (uint32_t)c.y;
    r.smap = c.map;
    r.patch = c.patch;
    r.pmap = c.pmap;
    r.score = lowest;
}

void build_luts(orc_job* j) {
    float sig2 = j->prm.cauchy_dispersion * j->prm.cauchy_dispersion;
    j->my_full.resize(65536);
    j->g_full.resize(65536);
    float a = j->alpha_s;
    for (int p = 0; p < 256; ++p)
        for (int q = 0; q < 256; ++q) {
            float c = metric_cauchy((uint8_t)p, (uint8_t)q, sig2);
            float l2 = metric_l2((uint8_t)p, (uint8_t)q);
            if (j->guided) {
                j->my_full[p * 256 + q] = (1.0f - a) * c;
                j->g_full[p * 256 + q] = a * l2;
            } else {
                j->my_full[p * 256 + q] = c;
                j->g_full[p * 256 + q] = 0.0f;
            }
        }
}

/* next_pyramid_level (ms.rs:687-700) */
bool recolor(orc_job* j) {
    for (uint32_t flat : j->resolved) {
        uint32_t m = j->cm[flat];
        if (m >= j->maps.size()) return set_err(j, "coord map id outside the example list");
        memcpy(&j->color[(size_t)flat * 4], j->maps[m]->at(j->cx[flat], j->cy[flat]), 4);
    }
    return true;
}

/* Session::run prelude: resolve_random_batch (ms.rs:427-445) then the
 * TreeGrid rebuild of resolve (ms.rs:747-779) */
bool init_run(orc_job* j) {
    if (j->prm.random_init != UINT64_MAX) {
        std::vector<const Img*> fine;
        for (auto& e : j->ex) fine.push_back(&e.back());
        for (uint64_t i = 0; i < j->prm.random_init; ++i) {
            uint32_t flat;
            if (!pick_unresolved(j, j->prm.seed + i, flat)) continue;
            uint32_t rmap, rx, ry;
            random_source(j->prm.seed + i + (uint64_t)flat, fine, rmap, rx, ry);
            write_pixel(j, flat, rx, ry, rmap, flat, rmap, fine);
            flush(j, flat, 0.0f, false);
        }
        j->locked += j->prm.random_init;
        if (j->locked > j->resolved.size()) return set_err(j, "random_init exceeds the pixels to resolve");
    }
    j->N0 = j->unresolved.size();
    /* every resolved pixel is re-flushed into the new grid with tiling ghosts */
    if (j->prm.tiling_mode)
        for (uint32_t flat : j->resolved) j->ghost[flat] = 1;
    uint32_t P = j->prm.backtrack_stages;
    j->ssum.assign(P + 1, 0.0);
    j->scnt.assign(P + 1, 0);
    j->sitems.assign(P + 1, 0);
    j->initialized = true;
    j->t = 0;
    j->stage_open = false;
    return true;
}

uint64_t stage_size(const orc_job* j, int s) {
    float v = std::pow(j->prm.backtrack_percent, (float)s) * (float)j->N0;
    if (!(v > 0.0f)) return 0;
    if (v >= 18446744073709551615.0f) return UINT64_MAX;
    return (uint64_t)v;
}

/* stage preamble (ms.rs:786-858) */
bool open_stage(orc_job* j) {
    int P = (int)j->prm.backtrack_stages;
    int s = P - j->t;
    j->level = j->t == 0 ? 0u : (uint32_t)std::min(j->t, P - 1);
    j->maps.clear();
    j->smask_f.clear();
    for (uint32_t fi : j->filt) {
        j->maps.push_back(&j->ex[fi][j->level]);
        j->smask_f.push_back(j->kind[fi] == 2 ? j->smask[fi].px.data() : nullptr);
    }
    j->gmaps.clear();
    if (j->guided) {
        for (auto& g : j->exg) j->gmaps.push_back(&g[j->level]);
        j->tgl = &j->tg[j->level];
    }
    if (j->level > 0 && !recolor(j)) return false;
    j->pss = (uint64_t)Pcg32::seed_from_u64(j->prm.seed + (uint64_t)s).next_u32();
    j->S = stage_size(j, s);
    j->redo = j->resolved.size() - j->locked;
    if (j->guided && s > 0) {
        float tr = (float)j->resolved.size();
        float b = j->prm.guide_alpha * (1.0f - (tr / (float)j->N0));
        j->alpha_s = (b * b) * b;
    } else {
        j->alpha_s = 0.0f;
    }
    build_luts(j);
    j->cursor = 0;
    j->stage_exhausted = false;
    j->stage_open = true;
    return true;
}

uint64_t batch_end(const orc_job* j, uint64_t i) {
    uint64_t redo_eff = std::min(j->redo, j->S);
    uint64_t seg_end = i < redo_eff ? redo_eff : j->S;
    uint64_t b;
    if (j->sched.px_per_batch > 0) {
        b = j->sched.px_per_batch;
    } else {
        uint64_t base = j->resolved.size() > j->locked ? j->resolved.size() - j->locked : 0;
        const uint32_t ns = (j->t == 0 && j->sched.first_new_shift) ? j->sched.first_new_shift : j->sched.new_shift;
        b = base >> (i < redo_eff ? j->sched.redo_shift : ns);
        if (b < 1) b = 1;
        if (j->sched.max_batch > 0 && b > j->sched.max_batch) b = j->sched.max_batch;
    }
    return std::min(i + b, seg_end);
}

/* pixels of items [b, e): redo items from the resolved list, new items
 * picked in order; returns false when the unresolved list runs dry */
size_t gather_items(orc_job* j, uint64_t b, uint64_t e, std::vector<uint32_t>& flats, std::vector<uint8_t>& is_new) {
    flats.clear();
    is_new.clear();
    for (uint64_t i = b; i < e; ++i) {
        if (i < j->redo) {
            flats.push_back(j->resolved[i + j->locked]);
            is_new.push_back(0);
        } else {
            uint32_t f;
            if (!pick_unresolved(j, j->pss + i, f)) {
                j->stage_exhausted = true;
                break;
            }
            flats.push_back(f);
            is_new.push_back(1);
        }
    }
    return flats.size();
}

bool run_batch(orc_job* j, uint64_t e) {
    std::vector<uint32_t> flats;
    std::vector<uint8_t> is_new;
    uint64_t b = j->cursor;
    size_t n = gather_items(j, b, e, flats, is_new);
    std::vector<Result> res(n);
    PointField f = field_of(j);
    if (j->scratch.size() < j->threads) j->scratch.resize(j->threads);
    parallel_for(n, j->threads, [&](size_t q, unsigned t) {
        evaluate(j, j->scratch[t], flats[q], is_new[q] != 0, j->pss + b + q, f, res[q]);
    });
    for (auto& s : j->scratch)
        if (s.error) return set_err(j, "map id outside the example list");
    /* commit in work order */
    for (size_t q = 0; q < n; ++q) {
        const Result& r = res[q];
        write_pixel(j, r.flat, r.sx, r.sy, r.smap, r.patch, r.pmap, j->maps);
        if (r.is_new) flush(j, r.flat, r.score, !r.fallback && j->prm.tiling_mode);
        /* per-stage statistics (not a reference quantity): items resolved by
         * pattern matching with a numeric cost (NaN weights, from a zero mean
         * neighbour distance, give NaN costs) */
        if (!r.fallback && r.score >= 0.0f) {
            j->ssum[j->t] += r.score;
            j->scnt[j->t] += 1;
        }
        j->sitems[j->t] += 1;
    }
    j->cursor = b + n;
    return true;
}

/* advance until (stage_t, item) or the end */
bool advance(orc_job* j, uint32_t stop_t, uint64_t stop_item) {
    if (!j->initialized && !init_run(j)) return false;
    int P = (int)j->prm.backtrack_stages;
    while (j->t <= P) {
        if (!j->stage_open && !open_stage(j)) return false;
        while (j->cursor < j->S && !j->stage_exhausted) {
            if ((uint32_t)j->t == stop_t && j->cursor >= stop_item) return true;
            uint64_t e = batch_end(j, j->cursor);
            if ((uint32_t)j->t == stop_t && e > stop_item) e = stop_item;
            if (!run_batch(j, e)) return false;
        }
        if ((uint32_t)j->t == stop_t) return true;
        j->t += 1;
        j->stage_open = false;
    }
    j->finished = true;
    return true;
}

/* ------------------------------------------------------------------ */
/* The reference's own threading (timing mode of the CPU baseline)     */
/* ------------------------------------------------------------------ */
/* ms.rs:815-1036: per stage, one worker while fewer than 1000 items are to
 * be redone, else N workers taking items from an atomic counter. Each item
 * reads the shared maps as they are and writes its result at once (the
 * reference's unsynchronised UnsafeCell state; here relaxed atomic word
 * accesses, so the restatement has no data race), and its k nearest come
 * from a TreeGrid (ms.rs:1313-1530): a grid of cells, each searched exactly
 * under a reader-writer lock (the reference's RwLock<RTree>), the query's
 * cell and its 8 neighbours with the reference's boundary pruning, rebuilt
 * with cells of sqrt(k W H / redo) * 2 + 1 pixels at the first fan-out.
 * Results depend on thread timing, as the reference's do; with one thread
 * the grid is one cell over the image (exact search, ties in the declared
 * (d^2, dy, dx) order) and the run equals one pixel per batch. */

namespace {

inline uint32_t ld32(const uint32_t* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
inline void st32(uint32_t* p, uint32_t v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }

struct LiveRead {
    const orc_job* j;
    NbrRec operator()(uint32_t f) const {
        return {ld32(&j->cx[f]), ld32(&j->cy[f]), ld32(&j->pid[f]), ld32(&j->pm[f])};
    }
};

/* One cell of the TreeGrid: its points bucketed by 8x8 pixels; exact k
 * nearest by rings of buckets around the query. */
struct GridCell {
    std::shared_mutex mu;
    int32_t x0 = 0, y0 = 0, nb = 1;  /* first pixel (grid coordinates) and buckets per side */
    std::vector<std::vector<uint64_t>> b;
    void init(int32_t gx0, int32_t gy0, int32_t size) {
        x0 = gx0;
        y0 = gy0;
        nb = (size + 7) / 8;
        b.assign((size_t)nb * nb, {});
    }
    void insert(int32_t gx, int32_t gy, int32_t x, int32_t y) {
        int32_t bx = std::min(nb - 1, std::max(0, (gx - x0) >> 3)), by = std::min(nb - 1, std::max(0, (gy - y0) >> 3));
        b[(size_t)by * nb + bx].push_back(((uint64_t)(uint32_t)x << 32) | (uint32_t)y);
    }
    /* the k nearest of (x, y) in this cell, ascending by (d^2, dy, dx), appended to out */
    void knn(int32_t gx, int32_t gy, int32_t x, int32_t y, uint32_t k, std::vector<KnnEntry>& out) const {
        std::vector<KnnEntry> best;
        auto worse = [](const KnnEntry& a, const KnnEntry& c) { return key_less(a.d2, a.dy, a.dx, c.d2, c.dy, c.dx); };
        const int32_t bx = std::min(nb - 1, std::max(0, (gx - x0) >> 3)), by = std::min(nb - 1, std::max(0, (gy - y0) >> 3));
        for (int32_t r = 0; r < nb; ++r) {
            if (best.size() >= k) {
                /* every point of ring r is at least (r - 1) * 8 pixels away */
                const int64_t lb = (int64_t)std::max(0, (r - 1) * 8);
                if (lb * lb > best.front().d2) break;
            }
            for (int32_t yy = by - r; yy <= by + r; ++yy) {
                if (yy < 0 || yy >= nb) continue;
                const bool edge = yy == by - r || yy == by + r;
                for (int32_t xx = bx - r; xx <= bx + r; xx += edge ? 1 : 2 * r) {
                    if (xx >= 0 && xx < nb)
                        for (uint64_t pt : b[(size_t)yy * nb + xx]) {
                            const int32_t px = (int32_t)(uint32_t)(pt >> 32), py = (int32_t)(uint32_t)pt;
                            const int64_t dx = px - x, dy = py - y;
                            const KnnEntry e{dx * dx + dy * dy, (int32_t)dy, (int32_t)dx};
                            if (best.size() < k) {
                                best.push_back(e);
                                std::push_heap(best.begin(), best.end(), worse);
                            } else if (worse(e, best.front())) {
                                std::pop_heap(best.begin(), best.end(), worse);
                                best.back() = e;
                                std::push_heap(best.begin(), best.end(), worse);
                            }
                        }
                    if (r == 0) break;
                }
            }
        }
        std::sort_heap(best.begin(), best.end(), worse);
        out.insert(out.end(), best.begin(), best.end());
    }
};

/* TreeGrid (ms.rs:1313-1530) */
struct TreeGridR {
    int32_t gw = 1, gh = 1, ox = 0, oy = 0, chunk = 1;
    std::vector<GridCell> cells;
    TreeGridR(uint32_t width, uint32_t height, uint32_t chunk_size, uint32_t offset_x, uint32_t offset_y)
        : cells(std::max<uint32_t>((width + chunk_size - 1) / chunk_size, 1) *
                std::max<uint32_t>((height + chunk_size - 1) / chunk_size, 1)) {
        gw = (int32_t)std::max<uint32_t>((width + chunk_size - 1) / chunk_size, 1);
        gh = (int32_t)std::max<uint32_t>((height + chunk_size - 1) / chunk_size, 1);
        ox = (int32_t)offset_x;
        oy = (int32_t)offset_y;
        chunk = (int32_t)chunk_size;
        for (int32_t cx = 0; cx < gw; ++cx)
            for (int32_t cy = 0; cy < gh; ++cy) cells[(size_t)cx * gh + cy].init(cx * chunk, cy * chunk, chunk);
    }
    GridCell* cell_of(int32_t x, int32_t y) {
        const int64_t cx = (int64_t)(uint32_t)(x + ox) / chunk, cy = (int64_t)(uint32_t)(y + oy) / chunk;
        if (cx >= gw || cy >= gh) return nullptr;  /* the reference would index out of its vector */
        return &cells[(size_t)cx * gh + (size_t)cy];
    }
    void insert(int32_t x, int32_t y) {
        GridCell* c = cell_of(x, y);
        if (!c) return;
        std::unique_lock<std::shared_mutex> lk(c->mu);
        c->insert(x + ox, y + oy, x, y);
    }
    void knn(uint32_t x, uint32_t y, uint32_t k, std::vector<int32_t>& nx, std::vector<int32_t>& ny) {
        nx.clear();
        ny.clear();
        const int32_t gx = (int32_t)x + ox, gy = (int32_t)y + oy;
        const int32_t cx = gx / chunk, cy = gy / chunk;
        const int64_t X = x, Y = y, C = chunk;
        struct Place {
            int32_t x, y;
            bool center;
            int64_t bx, by;
        };
        const Place places[9] = {
            {cx, cy, true, 0, 0},
            {cx + 1, cy, false, (cx + 1) * C - ox, Y},
            {cx - 1, cy, false, cx * C - ox, Y},
            {cx, cy - 1, false, X, cy * C - oy},
            {cx, cy + 1, false, X, (cy + 1) * C - oy},
            {cx + 1, cy + 1, false, (cx + 1) * C - ox, (cy + 1) * C - oy},
            {cx - 1, cy + 1, false, cx * C - ox, (cy + 1) * C - oy},
            {cx + 1, cy - 1, false, (cx + 1) * C - ox, cy * C - oy},
            {cx - 1, cy - 1, false, cx * C - ox, cy * C - oy},
        };
        std::vector<KnnEntry> tmp;
        tmp.reserve((size_t)k * 9);
        int64_t ub = INT64_MAX;
        for (const Place& pl : places) {
            if (pl.x < 0 || pl.x >= gw || pl.y < 0 || pl.y >= gh) continue;
            if (!pl.center) {
                const int64_t d2 = (X - pl.bx) * (X - pl.bx) + (Y - pl.by) * (Y - pl.by);
                if (d2 > ub) continue;
            }
            GridCell& c = cells[(size_t)pl.x * gh + pl.y];
            {
                std::shared_lock<std::shared_mutex> lk(c.mu);
                c.knn(gx, gy, (int32_t)x, (int32_t)y, k, tmp);
            }
            /* the reference's approximation of the k-th best distance */
            if (tmp.size() >= k && tmp.back().d2 < ub) ub = tmp.back().d2;
        }
        std::stable_sort(tmp.begin(), tmp.end(),
                         [](const KnnEntry& a, const KnnEntry& b) { return key_less(a.d2, a.dy, a.dx, b.d2, b.dy, b.dx); });
        for (size_t i = 0; i < tmp.size() && i < k; ++i) {
            nx.push_back((int32_t)x + tmp[i].dx);
            ny.push_back((int32_t)y + tmp[i].dy);
        }
    }
};

/* update (ms.rs:334-377): the maps at once; with update_list, flush_resolved
 * (ms.rs:296-331) into the grid and the worker's own list */
void live_update(orc_job* j, TreeGridR& g, std::vector<std::pair<uint32_t, float>>& mine, uint32_t flat, uint32_t sx,
                 uint32_t sy, uint32_t smap, uint32_t patch, uint32_t pmap, bool update_list, float score,
                 bool tiling) {
    st32(&j->cx[flat], sx);
    st32(&j->cy[flat], sy);
    st32(&j->cm[flat], smap);
    st32(&j->pid[flat], patch);
    st32(&j->pm[flat], pmap);
    uint32_t c;
    memcpy(&c, j->maps[smap]->at(sx, sy), 4);
    st32(reinterpret_cast<uint32_t*>(&j->color[(size_t)flat * 4]), c);
    if (!update_list) return;
    const int32_t x = (int32_t)(flat % j->W), y = (int32_t)(flat / j->W);
    g.insert(x, y);
    if (tiling) {
        const int32_t W = (int32_t)j->W, H = (int32_t)j->H;
        const int32_t xl = (int32_t)((float)W * 0.05f), xr = W - xl, yb = (int32_t)((float)H * 0.05f), yt = H - yb;
        if (x < xl) g.insert(x + W, y);
        else if (x > xr) g.insert(x - W, y);
        if (y < yb) g.insert(x, y + H);
        else if (y > yt) g.insert(x, y - H);
    }
    mine.push_back({flat, score});
}

bool run_reference_threads(orc_job* j, unsigned max_workers) {
    if (!j->initialized && !init_run(j)) return false;
    const bool tiling = j->prm.tiling_mode != 0;
    const uint32_t W = j->W, H = j->H, k = j->prm.nearest_neighbors;
    const uint32_t tw = (uint32_t)((float)W * std::fma(0.05f, 2.0f, 1.0f)) + 1;
    const uint32_t th = (uint32_t)((float)H * std::fma(0.05f, 2.0f, 1.0f)) + 1;
    const uint32_t offx = (uint32_t)((float)W * 0.05f) + 1, offy = (uint32_t)((float)H * 0.05f) + 1;
    /* the resolve prelude (ms.rs:747-779): one cell over the image */
    std::unique_ptr<TreeGridR> grid(new TreeGridR(tw, th, std::max(tw, th), offx, offy));
    std::vector<uint32_t> all_points;  /* everything inserted, to rebuild the grid at the fan-out */
    {
        std::vector<std::pair<uint32_t, float>> sink;
        std::vector<uint32_t> prev = j->resolved;
        std::vector<float> prev_s = j->rscore;
        j->resolved.clear();
        j->rscore.clear();
        for (size_t q = 0; q < prev.size(); ++q) {
            const int32_t x = (int32_t)(prev[q] % W), y = (int32_t)(prev[q] / W);
            grid->insert(x, y);
            if (tiling) {
                const int32_t xl = (int32_t)((float)W * 0.05f), xr = (int32_t)W - xl;
                const int32_t yb = (int32_t)((float)H * 0.05f), yt = (int32_t)H - yb;
                if (x < xl) grid->insert(x + (int32_t)W, y);
                else if (x > xr) grid->insert(x - (int32_t)W, y);
                if (y < yb) grid->insert(x, y + (int32_t)H);
                else if (y > yt) grid->insert(x, y - (int32_t)H);
            }
            j->resolved.push_back(prev[q]);
            j->rscore.push_back(prev_s[q]);
        }
    }
    bool fanned = false;
    std::mutex unresolved_mu;
    const int P = (int)j->prm.backtrack_stages;
    for (j->t = 0; j->t <= P; ++j->t) {
        if (!open_stage(j)) return false;
        const uint64_t redo_count = j->redo;
        const unsigned n_workers = redo_count < 1000 ? 1u : std::max(1u, max_workers);
        if (!fanned && n_workers > 1) {
            fanned = true;
            const uint32_t cell = (uint32_t)std::sqrt((double)(k * W * H / (uint32_t)redo_count)) * 2 + 1;
            std::unique_ptr<TreeGridR> ng(new TreeGridR(tw, th, cell, offx, offy));
            for (GridCell& c : grid->cells)
                for (auto& bk : c.b)
                    for (uint64_t pt : bk) ng->insert((int32_t)(uint32_t)(pt >> 32), (int32_t)(uint32_t)pt);
            grid = std::move(ng);
        }
        const uint64_t S = j->S;
        std::atomic<uint64_t> processed{0};
        std::vector<std::vector<std::pair<uint32_t, float>>> lists(n_workers);
        std::vector<double> ssum(n_workers, 0.0);
        std::vector<uint64_t> scnt(n_workers, 0), sitems(n_workers, 0);
        std::vector<Scratch> scr(n_workers);
        auto worker = [&](unsigned tid) {
            Scratch& s = scr[tid];
            for (;;) {
                const uint64_t i = processed.fetch_add(1, std::memory_order_relaxed);
                if (i >= S) break;
                const uint64_t loop_seed = j->pss + i;
                uint32_t flat;
                bool update_list;
                if (i < redo_count) {
                    update_list = false;
                    flat = j->resolved[i + j->locked];
                } else {
                    update_list = true;
                    std::lock_guard<std::mutex> lk(unresolved_mu);
                    if (!pick_unresolved(j, loop_seed, flat)) break;
                }
                const int32_t x = (int32_t)(flat % W), y = (int32_t)(flat / W);
                grid->knn((uint32_t)x, (uint32_t)y, k, s.nx, s.ny);
                if (!s.nx.empty()) {
                    weights_of(j, s, x, y);
                    find_candidates(j, s, x, y, loop_seed + 1, LiveRead{j});
                    const size_t kf = s.nx.size();
                    s.pat.resize(kf * 4);
                    for (size_t q = 0; q < kf; ++q) {
                        int32_t nx = s.nx[q], ny = s.ny[q];
                        if (tiling) {
                            nx = modulo(nx, (int32_t)W);
                            ny = modulo(ny, (int32_t)H);
                        }
                        const bool in = nx >= 0 && ny >= 0 && nx < (int32_t)W && ny < (int32_t)H;
                        uint32_t c = 0xff000000u;
                        if (in) c = ld32(reinterpret_cast<const uint32_t*>(&j->color[((size_t)ny * W + nx) * 4]));
                        memcpy(&s.pat[q * 4], &c, 4);
                    }
                    if (j->guided) pattern_of(*j->tgl, s, tiling, s.gpat);
                    size_t best;
                    float lowest;
                    find_best_match(j, s, best, lowest);
                    const Cand& c = s.cands[best];
                    live_update(j, *grid, lists[tid], flat, (uint32_t)c.x, (uint32_t)c.y, c.map, c.patch, c.pmap,
                                update_list, lowest, tiling);
                    if (lowest >= 0.0f) {
                        ssum[tid] += lowest;
                        scnt[tid] += 1;
                    }
                } else {
                    /* resolve_at_random (ms.rs:447-475) */
                    uint32_t rmap, rx, ry;
                    random_source(j->pss, j->maps, rmap, rx, ry);
                    live_update(j, *grid, lists[tid], flat, rx, ry, rmap, flat, rmap, true, 0.0f, false);
                }
                sitems[tid] += 1;
            }
        };
        if (n_workers == 1) {
            worker(0);
        } else {
            std::vector<std::thread> ths;
            for (unsigned t = 0; t < n_workers; ++t) ths.emplace_back(worker, t);
            for (auto& t : ths) t.join();
        }
        for (const Scratch& s : scr)
            if (s.error) return set_err(j, "map id outside the example list");
        /* the per-worker lists join the resolved list in worker order */
        for (unsigned t = 0; t < n_workers; ++t) {
            for (const auto& e : lists[t]) {
                j->mask[e.first] = 1;
                j->resolved.push_back(e.first);
                j->rscore.push_back(e.second);
            }
            j->ssum[j->t] += ssum[t];
            j->scnt[j->t] += scnt[t];
            j->sitems[j->t] += sitems[t];
        }
        j->stage_open = false;
    }
    j->finished = true;
    return true;
}

}  // namespace

Img to_img(const orc_image& im) {
    Img r;
    r.w = im.w;
    r.h = im.h;
    r.px.assign(im.rgba, im.rgba + (size_t)im.w * im.h * 4);
    return r;
}

/* transform_to_guide_map (utils.rs:101-116) without a resize: blur with
 * sigma 2 (the resampler at the same size), then grayscale (image 0.23
 * rgb_to_luma: 0.2126 R + 0.7152 G + 0.0722 B in f32, cast to u8) as RGBA. */
Img style_guide(const Img& src) {
    Img b;
    b.w = src.w;
    b.h = src.h;
    b.px.resize(src.px.size());
    resize_rgba(src.px.data(), src.w, src.h, src.w, src.h, 3, b.px.data());
    for (size_t i = 0; i < (size_t)b.w * b.h; ++i) {
        uint8_t* p = &b.px[i * 4];
        float l = 0.2126f * (float)p[0];
        l = l + 0.7152f * (float)p[1];
        l = l + 0.0722f * (float)p[2];
        const uint8_t v = (uint8_t)l;
        p[0] = p[1] = p[2] = v;
        p[3] = 255;
    }
    return b;
}

/* match_histograms / get_histogram / get_cdf (utils.rs:118-183) on the
 * first channel; the u8 arithmetic wraps as the release build's does. */
void match_histograms(Img& src, const Img& tgt) {
    auto cdf_of = [](const Img& im) {
        std::vector<uint32_t> hist(256, 0);
        for (size_t i = 0; i < (size_t)im.w * im.h; ++i) hist[im.px[i * 4]] += 1;
        std::vector<float> c(256, 0.0f);
        for (int i = 0; i < 256; ++i) c[i] = i ? c[i - 1] + (float)hist[i] : (float)hist[i];
        const float mx = c[255];
        for (float& v : c) v /= mx;
        return c;
    };
    const std::vector<float> tc = cdf_of(tgt), sc = cdf_of(src);
    for (size_t i = 0; i < (size_t)src.w * src.h; ++i) {
        const uint8_t v = src.px[i * 4];
        int pos = -1;
        for (int q = 0; q < 256; ++q)
            if (tc[q] > sc[v]) {
                pos = q;
                break;
            }
        const uint8_t n = (uint8_t)((pos >= 0 ? (uint8_t)pos : (uint8_t)(size_t)(uint8_t)(v + 1)) - 1u);
        uint8_t* p = &src.px[i * 4];
        p[0] = p[1] = p[2] = n;
        p[3] = 255;
    }
}

}  // namespace

/* ------------------------------------------------------------------ */
/* C API                                                               */
/* ------------------------------------------------------------------ */
extern "C" {

void orc_pcg_new_u32(uint64_t state, uint64_t stream, uint32_t n, uint32_t* out) {
    Pcg32 g = Pcg32::make(state, stream);
    for (uint32_t i = 0; i < n; ++i) out[i] = g.next_u32();
}
void orc_pcg_seed_u64(uint64_t seed, uint32_t n, uint64_t* out) {
    Pcg32 g = Pcg32::seed_from_u64(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = g.next_u64();
}
void orc_pcg_from_seed_u64(const uint8_t seed[16], uint32_t n, uint64_t* out) {
    Pcg32 g = Pcg32::from_seed(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = g.next_u64();
}
void orc_range_u32(uint64_t state, uint64_t stream, uint32_t lo, uint32_t hi, uint32_t n, uint32_t* out) {
    Pcg32 g = Pcg32::make(state, stream);
    for (uint32_t i = 0; i < n; ++i) out[i] = range_u32(g, lo, hi);
}
void orc_range_u64(uint64_t state, uint64_t stream, uint64_t lo, uint64_t hi, uint32_t n, uint64_t* out) {
    Pcg32 g = Pcg32::make(state, stream);
    for (uint32_t i = 0; i < n; ++i) out[i] = range_u64(g, lo, hi);
}
/* The neighbour weights' (float)exp(-1.0 * x) (ms.rs:1198-1203) with the host C
 * library's exp for the sweep arguments x_i (same formula as the device
 * test hook tsg_debug_exp_sweep), compared with `dev`: returns the number of
 * differing results; the first index that differs goes to *first_bad. */
uint64_t orc_exp_compare(uint64_t i0, uint64_t n, const float* dev, uint32_t threads, uint64_t* first_bad) {
    if (threads == 0) threads = 1;
    std::vector<uint64_t> bad(threads, 0), first(threads, UINT64_MAX);
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < threads; ++t)
        th.emplace_back([&, t] {
            const uint64_t lo = n * t / threads, hi = n * (t + 1) / threads;
            for (uint64_t r = lo; r < hi; ++r) {
                const uint64_t i = i0 + r;
                const double u = (double)((i * 0x9E3779B97F4A7C15ull) >> 11) * 0x1p-53;
                const double x = (i & 3u) ? u * 8.0 : 8.0 + u * 56.0;
                const float w = (float)std::exp(-1.0 * x);
                uint32_t a, b;
                memcpy(&a, &w, 4);
                memcpy(&b, &dev[r], 4);
                if (a != b) {
                    if (first[t] == UINT64_MAX) first[t] = i;
                    ++bad[t];
                }
            }
        });
    for (auto& x : th) x.join();
    uint64_t total = 0, f = UINT64_MAX;
    for (uint32_t t = 0; t < threads; ++t) {
        total += bad[t];
        f = std::min(f, first[t]);
    }
    if (first_bad) *first_bad = f;
    return total;
}

uint64_t orc_pick_index(uint64_t seed, uint64_t len) {
    Pcg32 g = Pcg32::seed_from_u64(seed);
    return range_u64(g, 0, len);
}

void orc_luts_full(float cauchy_dispersion, float alpha_s, int guided, float* my, float* gc) {
    float sig2 = cauchy_dispersion * cauchy_dispersion;
    for (int p = 0; p < 256; ++p)
        for (int q = 0; q < 256; ++q) {
            float c = metric_cauchy((uint8_t)p, (uint8_t)q, sig2);
            float l2 = metric_l2((uint8_t)p, (uint8_t)q);
            my[p * 256 + q] = guided ? (1.0f - alpha_s) * c : c;
            gc[p * 256 + q] = guided ? alpha_s * l2 : 0.0f;
        }
}

void orc_guide_map(const uint8_t* src, uint32_t w, uint32_t h, uint8_t* dst) {
    Img in;
    in.w = w;
    in.h = h;
    in.px.assign(src, src + (size_t)w * h * 4);
    const Img g = style_guide(in);
    memcpy(dst, g.px.data(), g.px.size());
}

void orc_match_histograms(uint8_t* src, uint32_t w, uint32_t h, const uint8_t* tgt, uint32_t tw, uint32_t th) {
    Img s, t;
    s.w = w;
    s.h = h;
    s.px.assign(src, src + (size_t)w * h * 4);
    t.w = tw;
    t.h = th;
    t.px.assign(tgt, tgt + (size_t)tw * th * 4);
    match_histograms(s, t);
    memcpy(src, s.px.data(), s.px.size());
}

int orc_resize(const uint8_t* src, uint32_t w, uint32_t h, uint32_t nw, uint32_t nh, int filter, uint8_t* dst) {
    if (!w || !h || !nw || !nh) return -1;
    resize_rgba(src, w, h, nw, nh, filter, dst);
    return 0;
}

int orc_pyramid(const uint8_t* src, uint32_t w, uint32_t h, uint32_t levels, uint8_t* out) {
    Img in;
    in.w = w;
    in.h = h;
    in.px.assign(src, src + (size_t)w * h * 4);
    std::vector<Img> pyr;
    if (!build_pyramid(in, levels == 0 ? 1 : levels, pyr)) return -1;
    for (size_t l = 0; l < pyr.size(); ++l) memcpy(out + l * (size_t)w * h * 4, pyr[l].px.data(), (size_t)w * h * 4);
    return 0;
}

orc_job* orc_job_new(const orc_params* prm, uint32_t n_examples, const orc_image* examples, const int* sample_kind,
                     const orc_image* sample_masks, const orc_image* example_guides, const orc_image* target_guide,
                     const orc_image* inpaint_mask, const orc_image* inpaint_color, uint32_t inpaint_index,
                     uint32_t out_w, uint32_t out_h) {
    orc_job* j = new orc_job();
    j->prm = *prm;
    j->W = out_w;
    j->H = out_h;
    j->levels = prm->backtrack_stages == 0 ? 1 : prm->backtrack_stages;
    if (!out_w || !out_h || n_examples == 0) {
        j->err = "invalid sizes";
        return j;
    }
    for (uint32_t i = 0; i < n_examples; ++i) {
        std::vector<Img> pyr;
        if (!build_pyramid(to_img(examples[i]), j->levels, pyr)) {
            j->err = "example too small for the pyramid";
            return j;
        }
        j->ex.push_back(std::move(pyr));
        int kd = sample_kind ? sample_kind[i] : 0;
        j->kind.push_back(kd);
        j->smask.push_back(kd == 2 ? to_img(sample_masks[i]) : Img());
        if (kd != 1) j->filt.push_back(i);
    }
    if (j->filt.empty()) {
        j->err = "no examples";
        return j;
    }
    if (target_guide) {
        j->guided = true;
        /* style transfer (session.rs:385-393, lib.rs:572-581): without example
         * guides the target guide becomes a guide map, and each example's
         * guide is its own guide map matched to the target's histogram */
        const Img tg = example_guides ? to_img(*target_guide) : style_guide(to_img(*target_guide));
        for (uint32_t i = 0; i < n_examples; ++i) {
            Img g;
            if (example_guides) {
                g = to_img(example_guides[i]);
            } else {
                g = style_guide(to_img(examples[i]));
                match_histograms(g, tg);
            }
            std::vector<Img> pyr;
            if (!build_pyramid(g, j->levels, pyr)) {
                j->err = "guide too small";
                return j;
            }
            j->exg.push_back(std::move(pyr));
        }
        if (!build_pyramid(tg, j->levels, j->tg)) {
            j->err = "target guide too small";
            return j;
        }
    }
    size_t N = (size_t)out_w * out_h;
    j->color.assign(N * 4, 0);
    j->cx.assign(N, 0);
    j->cy.assign(N, 0);
    j->cm.assign(N, 0);
    j->pid.assign(N, 0);
    j->pm.assign(N, 0);
    j->mask.assign(N, 0);
    j->ghost.assign(N, 0);
    if (inpaint_mask) {
        /* new_from_inpaint (ms.rs:236-293), inputs already at output size */
        memcpy(j->color.data(), inpaint_color->rgba, N * 4);
        for (size_t i = 0; i < N; ++i) {
            if (inpaint_mask->rgba[i * 4] < 255) {
                j->unresolved.push_back((uint32_t)i);
            } else {
                j->resolved.push_back((uint32_t)i);
                j->rscore.push_back(0.0f);
                j->mask[i] = 1;
                j->cx[i] = (uint32_t)(i % out_w);
                j->cy[i] = (uint32_t)(i / out_w);
                j->cm[i] = inpaint_index;
            }
        }
        j->locked = j->resolved.size();
    } else {
        j->unresolved.resize(N);
        for (size_t i = 0; i < N; ++i) j->unresolved[i] = (uint32_t)i;
    }
    return j;
}

void orc_job_free(orc_job* j) { delete j; }
const char* orc_job_error(orc_job* j) { return j->err.c_str(); }

int orc_job_run(orc_job* j, const orc_sched* s, uint32_t n_threads) {
    if (!j->err.empty()) return -1;
    j->sched = *s;
    j->threads = n_threads ? n_threads : 1;
    if (!advance(j, UINT32_MAX, 0)) return -1;
    return 0;
}

int orc_job_run_reference_threads(orc_job* j, uint32_t n_threads) {
    if (!j->err.empty()) return -1;
    if (j->initialized) return set_err(j, "the job has already started"), -1;
    if (!run_reference_threads(j, n_threads ? n_threads : 1)) return -1;
    return 0;
}

int orc_job_run_until(orc_job* j, const orc_sched* s, uint32_t n_threads, uint32_t stage_t, uint64_t item) {
    if (!j->err.empty()) return -1;
    j->sched = *s;
    j->threads = n_threads ? n_threads : 1;
    if (!advance(j, stage_t, item)) return -1;
    return 0;
}

void orc_job_get(orc_job* j, uint8_t* rgba, uint32_t* coord, uint32_t* ids) {
    size_t N = (size_t)j->W * j->H;
    if (rgba) memcpy(rgba, j->color.data(), N * 4);
    for (size_t i = 0; i < N; ++i) {
        if (coord) {
            coord[i * 3] = j->cx[i];
            coord[i * 3 + 1] = j->cy[i];
            coord[i * 3 + 2] = j->cm[i];
        }
        if (ids) {
            ids[i * 2] = j->pid[i];
            ids[i * 2 + 1] = j->pm[i];
        }
    }
}

uint64_t orc_job_resolved(orc_job* j, uint32_t* flat, float* score, uint64_t cap) {
    uint64_t n = std::min<uint64_t>(cap, j->resolved.size());
    for (uint64_t i = 0; i < n; ++i) {
        if (flat) flat[i] = j->resolved[i];
        if (score) score[i] = j->rscore[i];
    }
    return j->resolved.size();
}
uint64_t orc_job_locked(orc_job* j) { return j->locked; }

/* get_id_maps (ms.rs:603-630): Rgba([gen_range(0..255) of three seeded
 * streams, 255]); the seeds are u32 expressions of the id (wrapping, as the
 * release build computes them) widened to u64, except map * 200 in u64.
 * get_uncertainty_map (ms.rs:633-653): for each (pixel, score) of the
 * resolved list, n = (score.min(1.0) * 255.0) as u8 (saturating). */
void orc_job_debug_maps(orc_job* j, uint8_t* patch_ids, uint8_t* map_ids, uint8_t* uncertainty) {
    const size_t N = (size_t)j->W * j->H;
    auto draw = [](uint64_t seed) {
        Pcg32 g = Pcg32::seed_from_u64(seed);
        return (uint8_t)range_u8(g, 0, 255);
    };
    for (size_t i = 0; i < N; ++i) {
        const uint32_t patch = j->pid[i], map = j->pm[i];
        if (patch_ids) {
            uint8_t* o = patch_ids + 4 * i;
            o[0] = draw(patch);
            o[1] = draw((uint32_t)(patch * 5u + 21u));
            o[2] = draw((uint32_t)(patch / 4u + 12u));
            o[3] = 255;
        }
        if (map_ids) {
            uint8_t* o = map_ids + 4 * i;
            o[0] = draw((uint64_t)map * 200u);
            o[1] = draw((uint32_t)(map * 5u + 341u));
            o[2] = draw((uint32_t)(map * 1200u - 35412u));
            o[3] = 255;
        }
    }
    if (uncertainty) {
        memset(uncertainty, 0, N * 4);
        for (size_t q = 0; q < j->resolved.size(); ++q) {
            const float v = std::fmin(j->rscore[q], 1.0f) * 255.0f;
            const uint8_t n = !(v == v) ? 0 : v <= 0.0f ? 0 : v >= 255.0f ? 255 : (uint8_t)v;
            uint8_t* o = uncertainty + 4 * (size_t)j->resolved[q];
            o[0] = n;
            o[1] = (uint8_t)(255 - n);
            o[2] = 0;
            o[3] = 255;
        }
    }
}

uint32_t orc_job_stage_stats(orc_job* j, double* sums, uint64_t* counts, uint64_t* items) {
    for (size_t i = 0; i < j->ssum.size(); ++i) {
        if (sums) sums[i] = j->ssum[i];
        if (counts) counts[i] = j->scnt[i];
        if (items) items[i] = j->sitems[i];
    }
    return (uint32_t)j->ssum.size();
}

void orc_job_mask(orc_job* j, uint8_t* mask) { memcpy(mask, j->mask.data(), j->mask.size()); }

void orc_job_stage_info(orc_job* j, uint32_t* level, float* alpha_s, uint64_t* stage_seed, uint64_t* redo,
                        uint64_t* stage_size_out, uint64_t* cursor) {
    if (level) *level = j->level;
    if (alpha_s) *alpha_s = j->alpha_s;
    if (stage_seed) *stage_seed = j->pss;
    if (redo) *redo = j->redo;
    if (stage_size_out) *stage_size_out = j->S;
    if (cursor) *cursor = j->cursor;
}

void orc_job_stage_luts(orc_job* j, float* my256, float* g256) {
    for (int d = 0; d < 256; ++d) {
        if (my256) my256[d] = j->my_full[d];
        if (g256) g256[d] = j->g_full[d];
    }
}

uint64_t orc_job_describe(orc_job* j, uint64_t n, uint32_t* flat, uint32_t* nbr_count, int32_t* nbrs, float* weights,
                          uint32_t* cand_count, int32_t* cands, float* full_cost_out, uint32_t* best,
                          float* best_score) {
    if (!j->stage_open) return 0;
    uint64_t b = j->cursor;
    uint64_t e = std::min<uint64_t>(j->S, b + n);
    uint64_t redo_eff = std::min(j->redo, j->S);
    if (b < redo_eff && e > redo_eff) e = redo_eff;
    std::vector<uint32_t> flats;
    std::vector<uint8_t> is_new;
    size_t cnt = gather_items(j, b, e, flats, is_new);
    uint32_t k = j->prm.nearest_neighbors;
    size_t cmax = (size_t)k + (size_t)j->prm.random_sample_locations;
    PointField f = field_of(j);
    Scratch s;
    for (size_t q = 0; q < cnt; ++q) {
        Result r;
        evaluate(j, s, flats[q], is_new[q] != 0, j->pss + b + q, f, r);
        flat[q] = flats[q];
        size_t kf = r.fallback ? 0 : s.nx.size();
        nbr_count[q] = (uint32_t)kf;
        for (size_t a = 0; a < kf; ++a) {
            nbrs[(q * k + a) * 2] = s.nx[a];
            nbrs[(q * k + a) * 2 + 1] = s.ny[a];
            weights[q * k + a] = s.w[a];
        }
        size_t nc = r.fallback ? 0 : s.n_cands;
        cand_count[q] = (uint32_t)nc;
        for (size_t c = 0; c < nc; ++c) {
            const Cand& cd = s.cands[c];
            int32_t* o = &cands[(q * cmax + c) * 6];
            o[0] = cd.x;
            o[1] = cd.y;
            o[2] = (int32_t)cd.map;
            o[3] = (int32_t)cd.kind;
            o[4] = (int32_t)cd.patch;
            o[5] = (int32_t)cd.pmap;
            full_cost_out[q * cmax + c] = full_cost(j, s, cd);
        }
        if (r.fallback) {
            best[q] = UINT32_MAX;
            best_score[q] = 0.0f;
        } else {
            size_t bi;
            float lo;
            find_best_match(j, s, bi, lo);
            best[q] = (uint32_t)bi;
            best_score[q] = lo;
        }
    }
    return cnt;
}

void orc_knn(const uint8_t* mask, uint32_t w, uint32_t h, uint32_t n_q, const uint32_t* q_flat, uint32_t k,
             int32_t* nbrs, uint32_t* counts) {
    PointField f{mask, nullptr, (int32_t)w, (int32_t)h, 0, 0, 0, 0};
    std::vector<uint32_t> list;
    for (uint32_t i = 0; i < w * h; ++i)
        if (mask[i]) list.push_back(i);
    std::vector<int32_t> nx, ny;
    for (uint32_t q = 0; q < n_q; ++q) {
        knn_query(f, list, (int32_t)(q_flat[q] % w), (int32_t)(q_flat[q] / w), k, nx, ny);
        counts[q] = (uint32_t)nx.size();
        for (size_t a = 0; a < nx.size(); ++a) {
            nbrs[((size_t)q * k + a) * 2] = nx[a];
            nbrs[((size_t)q * k + a) * 2 + 1] = ny[a];
        }
    }
}

}  // extern "C"
