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
#include <math.h>

#include "kernels.h"

namespace tsg {

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerWG = 8;
constexpr int kKeyCap = 256;
constexpr int kLutCopies = 32;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) { return (1ull << lane) - 1ull; }

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
    uint32_t lo = __shfl((int)(uint32_t)v, src);
    uint32_t hi = __shfl((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
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

struct WaveLds {
    uint64_t keys[kKeyCap];
    double dist[kMaxK];
    int32_t nd[kMaxK];
    float w[kMaxK];
    uint32_t pat[kMaxK];
    uint32_t gpat[kMaxK];
    uint4 coh[kMaxK];
    double avg;
};

// ---- k-smallest selection over canonical keys, one wavefront ----
struct Sel {
    uint32_t count;
    uint64_t kth;  // current k-th smallest key once `count` reached k
};

__device__ void wave_sort(uint64_t* keys, int lane) {
    for (int size = 2; size <= kKeyCap; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < kKeyCap / 2; t += kWave) {
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

__device__ void sel_compact(uint64_t* keys, Sel& s, uint32_t k, int lane) {
    for (int t = lane; t < kKeyCap; t += kWave)
        if ((uint32_t)t >= s.count) keys[t] = ~0ull;
    wave_sync();
    wave_sort(keys, lane);
    s.count = s.count < k ? s.count : k;
    s.kth = s.count == k ? keys[k - 1] : ~0ull;
}

__device__ __forceinline__ void sel_push(uint64_t* keys, Sel& s, uint32_t k, bool valid, uint64_t key, int lane) {
    bool take = valid && key < s.kth;
    uint64_t m = __ballot(take);
    if (take) keys[s.count + __popcll(m & lanemask_lt(lane))] = key;
    s.count += (uint32_t)__popcll(m);
    if (s.count > kKeyCap - kWave) {
        wave_sync();
        sel_compact(keys, s, k, lane);
    }
}

__device__ __forceinline__ bool px_set(const uint64_t* mask, uint32_t tiles_x, int X, int Y) {
    uint64_t wd = mask[(uint32_t)(Y >> 3) * tiles_x + (uint32_t)(X >> 3)];
    return (wd >> (((Y & 7) << 3) | (X & 7))) & 1ull;
}

// Exhaustive scan of the resolved list (sparse phase).
__device__ void knn_list(const BatchArgs& a, WaveLds* L, int x, int y, uint32_t k, int lane, Sel& s) {
    for (uint32_t base = 0; base < a.n_res; base += kWave) {
        uint32_t idx = base + lane;
        bool v = idx < a.n_res;
        uint64_t key = ~0ull;
        if (v) {
            uint32_t f = a.rlist[idx];
            int X = (int)(f % a.W), Y = (int)(f / a.W);
            key = make_key(X - x, Y - y);
        }
        sel_push(L->keys, s, k, v, key, lane);
    }
}

// Ring search over 8x8 mask tiles, nearest tile rings first; stops once the
// k-th best distance is below the distance to any unvisited tile.
__device__ void knn_rings(const BatchArgs& a, WaveLds* L, int x, int y, uint32_t k, int lane, Sel& s) {
    const int tx = x >> 3, ty = y >> 3;
    const int TX = (int)a.tiles_x, TY = (int)a.tiles_y;
    for (int r = 0;; ++r) {
        int nt = r == 0 ? 1 : 8 * r;
        for (int tb = 0; tb < nt; tb += kWave) {
            int t = tb + lane;
            uint64_t word = 0;
            int gx = 0, gy = 0;
            if (t < nt) {
                int ox = 0, oy = 0;
                if (r > 0) {
                    int side = t / (2 * r), o = t - side * (2 * r);
                    if (side == 0) {
                        ox = -r + o;
                        oy = -r;
                    } else if (side == 1) {
                        ox = r;
                        oy = -r + o;
                    } else if (side == 2) {
                        ox = r - o;
                        oy = r;
                    } else {
                        ox = -r;
                        oy = r - o;
                    }
                }
                gx = tx + ox;
                gy = ty + oy;
                if (gx >= 0 && gy >= 0 && gx < TX && gy < TY) word = a.mask[(uint32_t)gy * a.tiles_x + (uint32_t)gx];
            }
            uint64_t nz = __ballot(word != 0);
            while (nz) {
                int l = __ffsll((long long)nz) - 1;
                nz &= nz - 1;
                uint64_t wd = shfl_u64(word, l);
                int bx = __shfl(gx, l), by = __shfl(gy, l);
                bool b = (wd >> lane) & 1ull;
                int X = bx * 8 + (lane & 7), Y = by * 8 + (lane >> 3);
                sel_push(L->keys, s, k, b, make_key(X - x, Y - y), lane);
            }
        }
        bool all = (tx - r <= 0) && (ty - r <= 0) && (tx + r >= TX - 1) && (ty + r >= TY - 1);
        if (all) break;
        if (s.count >= k) {
            wave_sync();
            sel_compact(L->keys, s, k, lane);
            int64_t D = INT64_MAX, e;
            if (tx - r > 0 && (e = x - (tx - r) * 8 + 1) < D) D = e;
            if (tx + r < TX - 1 && (e = (tx + r + 1) * 8 - x) < D) D = e;
            if (ty - r > 0 && (e = y - (ty - r) * 8 + 1) < D) D = e;
            if (ty + r < TY - 1 && (e = (ty + r + 1) * 8 - y) < D) D = e;
            if ((int64_t)(s.kth >> 32) < D * D) break;
        }
    }
}

// Ring-ordered offset table scan: the first k resolved pixels in table
// order are the k nearest (the table lists every offset within its radius
// in canonical order). Returns the number found (<= k).
__device__ uint32_t knn_table(const BatchArgs& a, WaveLds* L, int x, int y, uint32_t k, int lane) {
    uint32_t found = 0;
    for (uint32_t base = 0; base < a.offtab_n; base += kWave) {
        uint32_t t = base + lane;
        bool hit = false;
        int32_t o = 0;
        if (t < a.offtab_n) {
            o = a.offtab[t];
            int X = x + packed_dx(o), Y = y + packed_dy(o);
            hit = X >= 0 && Y >= 0 && X < (int)a.W && Y < (int)a.H && px_set(a.mask, a.tiles_x, X, Y);
        }
        uint64_t m = __ballot(hit);
        uint32_t pre = (uint32_t)__popcll(m & lanemask_lt(lane));
        if (hit && found + pre < k) L->nd[found + pre] = o;
        found += (uint32_t)__popcll(m);
        if (found >= k) break;
    }
    return found < k ? found : k;
}

__device__ __forceinline__ bool coord_ok(const MapDesc& md, int x, int y) {
    if (x < 0 || y < 0 || (uint32_t)x >= md.w || (uint32_t)y >= md.h) return false;
    return md.smask == nullptr || md.smask[(uint32_t)y * md.w + (uint32_t)x] != 0;
}

__device__ __forceinline__ float lut4(const float* lut, uint32_t q, uint32_t p, int slot) {
    // four channel terms in R, G, B, A order into one accumulator
    int d0 = abs((int)(q & 255u) - (int)(p & 255u));
    int d1 = abs((int)((q >> 8) & 255u) - (int)((p >> 8) & 255u));
    int d2 = abs((int)((q >> 16) & 255u) - (int)((p >> 16) & 255u));
    int d3 = abs((int)(q >> 24) - (int)(p >> 24));
    float s = lut[d0 * kLutCopies + slot];
    s = s + lut[d1 * kLutCopies + slot];
    s = s + lut[d2 * kLutCopies + slot];
    s = s + lut[d3 * kLutCopies + slot];
    return s;
}

template <bool GUIDED>
__global__ __launch_bounds__(kWave* kWavesPerWG) void k_resolve(BatchArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* lut = reinterpret_cast<float*>(smem);
    float* glut = lut + 256 * kLutCopies;
    WaveLds* wls = reinterpret_cast<WaveLds*>(smem + (GUIDED ? 2 : 1) * 256 * kLutCopies * sizeof(float));

    for (int t = threadIdx.x; t < 256 * kLutCopies; t += blockDim.x) {
        lut[t] = a.lut[t / kLutCopies];
        if (GUIDED) glut[t] = a.lut[256 + t / kLutCopies];
    }
    __syncthreads();

    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int slot = lane % kLutCopies;
    const uint32_t item = a.item0 + blockIdx.x * kWavesPerWG + wave;
    if (item >= a.item1) return;
    WaveLds* L = wls + wave;
    const uint32_t rel = item - a.item0;
    const uint32_t flat = a.work[item];
    const int x = (int)(flat % a.W), y = (int)(flat / a.W);
    const uint32_t k = a.k;

    // ---- 1. k nearest resolved neighbours, canonical order ----
    uint32_t kf;
    if (a.ext_nbrs) {
        kf = a.ext_nbr_count[rel];
        if (lane < (int)kf) L->nd[lane] = a.ext_nbrs[(size_t)rel * k + lane];
    } else {
        uint32_t found = 0;
        bool use_sel = a.n_res <= kBruteMax;
        if (!use_sel) {
            found = knn_table(a, L, x, y, k, lane);
            use_sel = found < k && found < a.n_res;
        }
        if (use_sel) {
            Sel s{0u, ~0ull};
            if (a.n_res <= kBruteMax) knn_list(a, L, x, y, k, lane, s);
            else knn_rings(a, L, x, y, k, lane, s);
            wave_sync();
            sel_compact(L->keys, s, k, lane);
            found = s.count;
            if (lane < (int)found) L->nd[lane] = key_to_packed(L->keys[lane]);
        }
        kf = found;
    }
    wave_sync();
    if (a.dbg_nbrs) {
        if (lane < (int)kf) a.dbg_nbrs[(size_t)rel * k + lane] = L->nd[lane];
        if (lane == 0) a.dbg_nbr_count[rel] = kf;
    }
    if (a.knn_only) return;

    if (kf == 0) {
        // no resolved neighbour: resolve_at_random with the stage seed (ms.rs:1003-1009)
        if (lane == 0) {
            Pcg32 g0 = Pcg32::seed_from_u64(a.stage_seed);
            uint32_t rm = (uint32_t)range_u64(g0, 0, a.n_maps);
            Pcg32 g1 = Pcg32::seed_from_u64(a.stage_seed);
            uint32_t rx = range_u32(g1, 0, a.maps[rm].w);
            Pcg32 g2 = Pcg32::seed_from_u64(a.stage_seed);
            uint32_t ry = range_u32(g2, 0, a.maps[rm].h);
            a.result[rel] = make_uint4(rx | (ry << 16), rm | (rm << 16) | kFallbackBit, flat, __float_as_uint(0.0f));
            if (a.item_score) a.item_score[item] = -1.0f;
            if (a.dbg_best) a.dbg_best[rel] = 0xffffffffu;
        }
        return;
    }

    // ---- 2. neighbour state: own pattern, distances, coherence candidates ----
    bool coh_valid = false;
    uint4 coh = make_uint4(0, 0, 0, 0);
    if (lane < (int)kf) {
        int32_t o = L->nd[lane];
        int nx = x + packed_dx(o), ny = y + packed_dy(o);
        uint32_t nflat = (uint32_t)ny * a.W + (uint32_t)nx;
        uint4 st = a.state[nflat];
        L->pat[lane] = st.x;
        if (GUIDED) L->gpat[lane] = a.tg[nflat];
        double dimx = (double)a.W, dimy = (double)a.H;
        double x1 = (double)nx / dimx, y1 = (double)ny / dimy;
        double x2 = (double)x / dimx, y2 = (double)y / dimy;
        L->dist[lane] = fma(x1 - x2, x1 - x2, (y1 - y2) * (y1 - y2));
        if (!a.ext_cands) {
            int cx = (int)(st.y & 0xffffu) - packed_dx(o);
            int cy = (int)(st.y >> 16) - packed_dy(o);
            uint32_t pmap = st.z >> 16;
            coh_valid = pmap < a.n_maps && coord_ok(a.maps[pmap], cx, cy);
            coh = make_uint4((uint32_t)cx | ((uint32_t)cy << 16), pmap, st.w, 0u);
        }
    }
    uint64_t cm = __ballot(coh_valid);
    if (coh_valid) L->coh[__popcll(cm & lanemask_lt(lane))] = coh;
    const uint32_t n_coh = (uint32_t)__popcll(cm);
    wave_sync();
    if (lane == 0) {
        double sum = 0.0;
        for (uint32_t j = 0; j < kf; ++j) {
            double d = L->dist[j];
            sum = sum + d;
            sum = sum + d;
            sum = sum + d;
            sum = sum + d;
        }
        L->avg = sum / (double)(4u * kf);
    }
    wave_sync();
    if (lane < (int)kf) L->w[lane] = (float)exp(-1.0 * (L->dist[lane] / L->avg));
    wave_sync();

    // ---- 3. candidate costs, one lane per candidate ----
    const uint32_t n_c = a.ext_cands ? a.ext_cand_count[rel] : n_coh + a.m;
    float best = FLT_MAX;
    uint32_t bidx = 0;
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
                uint2 r = a.rcands[(size_t)(item - a.rc_base) * a.m + (c - n_coh)];
                cxy = r.x;
                cmap = r.y;
                sgn = -1;
            }
            const MapDesc md = a.maps[cmap];
            const int cx = (int)(cxy & 0xffffu), cy = (int)(cxy >> 16);
            float score = 0.0f, lastgood = -INFINITY;
            for (uint32_t j = 0; j < kf; ++j) {
                const int32_t o = L->nd[j];
                const int X2 = cx + sgn * packed_dx(o), Y2 = cy + sgn * packed_dy(o);
                const bool in = (uint32_t)X2 < md.w && (uint32_t)Y2 < md.h;
                const uint32_t idx = (uint32_t)Y2 * md.w + (uint32_t)X2;
                const uint32_t p = in ? md.px[idx] : kOutsidePx;
                float s = lut4(lut, L->pat[j], p, slot);
                if (GUIDED) {
                    const bool gin = (uint32_t)X2 < md.gw && (uint32_t)Y2 < md.gh;
                    const uint32_t gp = gin ? md.gpx[(uint32_t)Y2 * md.gw + (uint32_t)X2] : kOutsidePx;
                    const uint32_t q = L->gpat[j];
                    s = s + glut[abs((int)(q & 255u) - (int)(gp & 255u)) * kLutCopies + slot];
                    s = s + glut[abs((int)((q >> 8) & 255u) - (int)((gp >> 8) & 255u)) * kLutCopies + slot];
                    s = s + glut[abs((int)((q >> 16) & 255u) - (int)((gp >> 16) & 255u)) * kLutCopies + slot];
                    s = s + glut[abs((int)(q >> 24) - (int)(gp >> 24)) * kLutCopies + slot];
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

    // ---- 4. winner ----
    if (lane == 0) {
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
            uint2 r = a.rcands[(size_t)(item - a.rc_base) * a.m + (bidx - n_coh)];
            cxy = r.x;
            cmap = r.y;
            patch = (cxy >> 16) * a.maps[cmap].w + (cxy & 0xffffu);
        }
        a.result[rel] = make_uint4(cxy, cmap | (cmap << 16), patch, __float_as_uint(best));
        if (a.item_score) a.item_score[item] = best;
        if (a.dbg_best) a.dbg_best[rel] = bidx;
    }
}

// Random candidates of items [item0, item1): Pcg32::seed_from_u64(stage
// seed + i + 1), m draws of (map, then x and y until the sample mask accepts).
__global__ void k_rand(RandArgs a) {
    uint32_t item = a.item0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= a.item1) return;
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

__global__ void k_commit(CommitArgs a) {
    uint32_t item = a.item0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= a.item1) return;
    uint32_t p = a.work[item];
    uint4 r = a.result[item - a.item0];
    uint32_t cmap = r.y & 0xffffu;
    const MapDesc md = a.maps[cmap];
    uint32_t sx = r.x & 0xffffu, sy = r.x >> 16;
    uint32_t color = md.px[sy * md.w + sx];
    a.state[p] = make_uint4(color, r.x, r.y & ~kFallbackBit, r.z);
    if (item >= a.first_new) {
        uint32_t px = p % a.W, py = p / a.W;
        atomicOr((unsigned long long*)&a.mask[(py >> 3) * a.tiles_x + (px >> 3)],
                 1ull << (((py & 7u) << 3) | (px & 7u)));
        a.score[p] = __uint_as_float(r.w);
    }
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

}  // namespace

hipError_t configure_kernels() {
    hipError_t e = hipFuncSetAttribute((const void*)k_resolve<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)resolve_lds_bytes(true));
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)k_resolve<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)resolve_lds_bytes(false));
}

hipError_t launch_stats(const StatsArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    uint32_t blocks = (a.n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_stats, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

size_t resolve_lds_bytes(bool guided) {
    return (guided ? 2 : 1) * 256 * kLutCopies * sizeof(float) + kWavesPerWG * sizeof(WaveLds);
}

hipError_t launch_resolve(const BatchArgs& a, bool guided, hipStream_t s) {
    uint32_t n = a.item1 - a.item0;
    if (n == 0) return hipSuccess;
    dim3 grid((n + kWavesPerWG - 1) / kWavesPerWG), block(kWave * kWavesPerWG);
    size_t lds = resolve_lds_bytes(guided);
    if (guided) {
        hipLaunchKernelGGL(k_resolve<true>, grid, block, lds, s, a);
    } else {
        hipLaunchKernelGGL(k_resolve<false>, grid, block, lds, s, a);
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
