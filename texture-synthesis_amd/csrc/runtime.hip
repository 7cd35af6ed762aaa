// Host runtime of the patch-search engine: contexts, jobs (device-resident
// pyramids and state), the backtracking stage schedule and the C ABI.
//
// Session::run (lib/src/session.rs:37-66) = random init (ms.rs:427-445) +
// Generator::resolve (ms.rs:702-1051). Per stage the host derives the level,
// stage seed, stage size, adaptive alpha and 256-entry cost tables
// (ms.rs:786-858), picks the new pixels in the reference's order
// (pick_random_unresolved, ms.rs:380-389), and enqueues Jacobi batches of
// {k_resolve, k_commit} on one stream without waiting on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "../../include/tsgpu.h"
#include "exp_dd.h"
#include "kernels.h"

using namespace tsg;

// Device scratch of the entry points that allocate nothing
// (tsg_knn_batch_noalloc): staging of inputs and outputs and every
// intermediate buffer, carved per call. One copy per device, held until the
// process ends. Host code never uses this name as a device address: each
// context resolves the address on its device once (hipGetSymbolAddress in
// tsg_ctx_create) and passes that pointer to copies and kernels. All calls on
// one device share the array, so calls that overlap on one device race.
constexpr size_t kScratchBytes = 32ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            int c_ = e_ == hipErrorOutOfMemory ? TSG_ERR_OUT_OF_MEMORY : TSG_ERR_DEVICE;           \
            return fail(c_, std::string(#expr) + ": " + hipGetErrorString(e_));                    \
        }                                                                                          \
    } while (0)

// Device blocks of destroyed jobs, per device, kept for the next jobs'
// allocations: a job allocates about a hundred buffers (pyramids, state,
// stage-ahead sets) and hipMalloc / hipFree cost tens of microseconds each.
// Only releases made while g_pool_release is set (job destruction, when the
// job has no device work in flight) enter the cache; any other release may
// race queued kernels and goes to hipFree, which waits for them. At most
// kPoolCap bytes per device stay cached; a block serves requests of half its
// size and more.
constexpr size_t kPoolCap = 4ull << 30;
constexpr int kPoolDevices = 16;
thread_local bool g_pool_release = false;

struct DevPool {
    std::mutex mu;
    std::multimap<size_t, void*> free;  // bytes -> block
    size_t cached = 0;
};
DevPool g_pool[kPoolDevices];

DevPool* pool_here() {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= kPoolDevices) return nullptr;
    return &g_pool[d];
}

// Test hook (tsg_debug_fail_allocs): the next this-many device allocations
// of the library fail as if the device were out of memory.
std::atomic<int64_t> g_fail_allocs{0};

// A cached block of at least `bytes` (and at most twice that), or a new one.
hipError_t pool_get(size_t bytes, void** p, size_t* got) {
    if (g_fail_allocs.load(std::memory_order_relaxed) > 0 && g_fail_allocs.fetch_sub(1) > 0)
        return hipErrorOutOfMemory;
    if (DevPool* pl = pool_here()) {
        std::lock_guard<std::mutex> lk(pl->mu);
        auto it = pl->free.lower_bound(bytes);
        if (it != pl->free.end() && it->first <= 2 * bytes) {
            *p = it->second;
            *got = it->first;
            pl->cached -= it->first;
            pl->free.erase(it);
            return hipSuccess;
        }
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {
        // give the cache back and retry once
        if (DevPool* pl = pool_here()) {
            std::lock_guard<std::mutex> lk(pl->mu);
            for (auto& kv : pl->free) (void)hipFree(kv.second);
            pl->free.clear();
            pl->cached = 0;
        }
        e = hipMalloc(p, bytes);
    }
    if (e == hipSuccess) *got = bytes;
    return e;
}

void pool_put(void* p, size_t bytes) {
    if (g_pool_release) {
        if (DevPool* pl = pool_here()) {
            std::lock_guard<std::mutex> lk(pl->mu);
            if (pl->cached + bytes <= kPoolCap) {
                pl->free.emplace(bytes, p);
                pl->cached += bytes;
                return;
            }
        }
    }
    (void)hipFree(p);
}

// Owning device allocation.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    size_t bytes = 0;  // size of the block behind p
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) pool_put(p, bytes);
        p = nullptr;
        n = 0;
        bytes = 0;
    }
    hipError_t alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        void* q = nullptr;
        size_t got = 0;
        hipError_t e = pool_get(count * sizeof(T), &q, &got);
        if (e == hipSuccess) {
            p = static_cast<T*>(q);
            n = count;
            bytes = got;
        }
        return e;
    }
    hipError_t reserve(size_t count) { return count <= n ? hipSuccess : alloc(count); }
};

// ---- image 0.23 separable resampling weights (imageops/sample.rs) ----
struct Filter {
    float (*kernel)(float);
    float support;
};
float kern_triangle(float x) {
    float a = std::fabs(x);
    return a < 1.0f ? 1.0f - a : 0.0f;
}
float kern_catmullrom(float x) {
    // bc_cubic_spline with B = 0, C = 0.5
    const float b = 0.0f, c = 0.5f;
    float a = std::fabs(x), k;
    if (a < 1.0f)
        k = (12.0f - 9.0f * b - 6.0f * c) * (a * a * a) + (-18.0f + 12.0f * b + 6.0f * c) * (a * a) + (6.0f - 2.0f * b);
    else if (a < 2.0f)
        k = (-b - 6.0f * c) * (a * a * a) + (6.0f * b + 30.0f * c) * (a * a) + (-12.0f * b - 48.0f * c) * a +
            (8.0f * b + 24.0f * c);
    else
        k = 0.0f;
    return k / 6.0f;
}
// imageops::blur(image, 2.0): gaussian(x, sigma) with support 2 sigma
float kern_gaussian_blur2(float x) {
    const float r = 2.0f;
    float c = 1.0f / (std::sqrt(2.0f * 3.14159265358979323846f) * r);
    return c * std::exp(-(x * x) / (2.0f * (r * r)));
}
float kern_gaussian(float x) {
    // gaussian(x, r = 0.5)
    const float r = 0.5f;
    float c = 1.0f / (std::sqrt(2.0f * 3.14159265358979323846f) * r);
    return c * std::exp(-(x * x) / (2.0f * (r * r)));
}

struct Taps {
    std::vector<uint32_t> left, cnt;
    std::vector<float> ws;
    uint32_t stride = 0;
};

Taps make_taps(uint32_t in_len, uint32_t out_len, const Filter& f) {
    Taps t;
    float ratio = (float)in_len / (float)out_len;
    float sratio = ratio < 1.0f ? 1.0f : ratio;
    float support = f.support * sratio;
    std::vector<std::vector<float>> tmp(out_len);
    t.left.resize(out_len);
    t.cnt.resize(out_len);
    for (uint32_t o = 0; o < out_len; ++o) {
        float in = ((float)o + 0.5f) * ratio;
        int64_t l = (int64_t)std::floor(in - support);
        l = std::min<int64_t>(std::max<int64_t>(l, 0), (int64_t)in_len - 1);
        int64_t r = (int64_t)std::ceil(in + support);
        r = std::min<int64_t>(std::max<int64_t>(r, l + 1), (int64_t)in_len);
        in = in - 0.5f;
        float sum = 0.0f;
        for (int64_t i = l; i < r; ++i) {
            float w = f.kernel(((float)i - in) / sratio);
            tmp[o].push_back(w);
            sum += w;
        }
        for (float& w : tmp[o]) w /= sum;
        t.left[o] = (uint32_t)l;
        t.cnt[o] = (uint32_t)(r - l);
        t.stride = std::max(t.stride, t.cnt[o]);
    }
    t.ws.assign((size_t)out_len * t.stride, 0.0f);
    for (uint32_t o = 0; o < out_len; ++o)
        for (uint32_t i = 0; i < t.cnt[o]; ++i) t.ws[(size_t)o * t.stride + i] = tmp[o][i];
    return t;
}

struct DevTaps {
    DBuf<uint32_t> left, cnt;
    DBuf<float> ws;
    uint32_t stride = 0;
    hipError_t upload(const Taps& t, hipStream_t s) {
        hipError_t e;
        if ((e = left.alloc(t.left.size())) != hipSuccess) return e;
        if ((e = cnt.alloc(t.cnt.size())) != hipSuccess) return e;
        if ((e = ws.alloc(t.ws.size())) != hipSuccess) return e;
        stride = t.stride;
        if ((e = hipMemcpyAsync(left.p, t.left.data(), t.left.size() * 4, hipMemcpyHostToDevice, s)) != hipSuccess)
            return e;
        if ((e = hipMemcpyAsync(cnt.p, t.cnt.data(), t.cnt.size() * 4, hipMemcpyHostToDevice, s)) != hipSuccess)
            return e;
        return hipMemcpyAsync(ws.p, t.ws.data(), t.ws.size() * 4, hipMemcpyHostToDevice, s);
    }
};

// imageops::resize(img, nw, nh, filter) on the device: vertical then
// horizontal pass (f32 in between, rounded u8 out).
int device_resize(const uint32_t* src, uint32_t w, uint32_t h, uint32_t* dst, uint32_t nw, uint32_t nh,
                  const Filter& f, hipStream_t s) {
    Taps tv = make_taps(h, nh, f), th = make_taps(w, nw, f);
    DevTaps dv, dh;
    DBuf<float4> tmp;
    HIP_TRY(dv.upload(tv, s));
    HIP_TRY(dh.upload(th, s));
    HIP_TRY(tmp.alloc((size_t)w * nh));
    HIP_TRY(launch_resample_v(src, w, h, tmp.p, nh, dv.left.p, dv.cnt.p, dv.ws.p, dv.stride, s));
    HIP_TRY(launch_resample_h(tmp.p, w, nh, dst, nw, dh.left.p, dh.cnt.p, dh.ws.p, dh.stride, s));
    HIP_TRY(hipStreamSynchronize(s));  // taps and tmp die here
    return TSG_OK;
}

// transform_to_guide_map (utils.rs:101-116) without a resize: blur with
// sigma 2 on the device (the resampler at the same size), then grayscale on
// the host (image 0.23 rgb_to_luma: 0.2126 R + 0.7152 G + 0.0722 B in f32,
// cast to u8), as RGBA.
int style_guide(hipStream_t s, const tsg_image& im, std::vector<uint8_t>& out) {
    const size_t n = (size_t)im.w * im.h;
    DBuf<uint32_t> src, dst;
    HIP_TRY(src.alloc(n));
    HIP_TRY(dst.alloc(n));
    HIP_TRY(hipMemcpyAsync(src.p, im.rgba, n * 4, hipMemcpyHostToDevice, s));
    const Filter blur{kern_gaussian_blur2, 4.0f};
    int rc = device_resize(src.p, im.w, im.h, dst.p, im.w, im.h, blur, s);
    if (rc) return rc;
    out.resize(n * 4);
    HIP_TRY(hipMemcpy(out.data(), dst.p, n * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) {
        uint8_t* p = &out[i * 4];
        float l = 0.2126f * (float)p[0];
        l = l + 0.7152f * (float)p[1];
        l = l + 0.0722f * (float)p[2];
        const uint8_t v = (uint8_t)l;
        p[0] = p[1] = p[2] = v;
        p[3] = 255;
    }
    return TSG_OK;
}

// match_histograms / get_histogram / get_cdf (utils.rs:118-183) on the first
// channel: each value moves to the position before the first target CDF entry
// above its own CDF; the u8 arithmetic wraps as the release build's does.
void match_histograms(std::vector<uint8_t>& src, const std::vector<uint8_t>& tgt) {
    auto cdf_of = [](const std::vector<uint8_t>& im) {
        std::vector<uint32_t> hist(256, 0);
        for (size_t i = 0; i < im.size(); i += 4) hist[im[i]] += 1;
        std::vector<float> c(256, 0.0f);
        for (int i = 0; i < 256; ++i) c[i] = i ? c[i - 1] + (float)hist[i] : (float)hist[i];
        const float mx = c[255];
        for (float& v : c) v /= mx;
        return c;
    };
    const std::vector<float> tc = cdf_of(tgt), sc = cdf_of(src);
    uint8_t map[256];
    for (int v = 0; v < 256; ++v) {
        int pos = -1;
        for (int q = 0; q < 256; ++q)
            if (tc[q] > sc[v]) {
                pos = q;
                break;
            }
        map[v] = (uint8_t)((pos >= 0 ? (uint8_t)pos : (uint8_t)(v + 1)) - 1u);
    }
    for (size_t i = 0; i < src.size(); i += 4) {
        const uint8_t n = map[src[i]];
        src[i] = src[i + 1] = src[i + 2] = n;
        src[i + 3] = 255;
    }
}

// ImagePyramid::build_gaussian (img_pyramid.rs:20-37): `levels` full-size
// levels at out, coarsest first; the last is the input itself.
int device_pyramid(const uint8_t* host_rgba, uint32_t w, uint32_t h, uint32_t levels, uint32_t* out,
                   hipStream_t s) {
    const Filter g{kern_gaussian, 3.0f};
    size_t npx = (size_t)w * h;
    uint32_t* last = out + (size_t)(levels - 1) * npx;
    HIP_TRY(hipMemcpyAsync(last, host_rgba, npx * 4, hipMemcpyHostToDevice, s));
    DBuf<uint32_t> small;
    for (uint32_t lv = 0; lv + 1 < levels; ++lv) {
        uint32_t i = levels - 1 - lv;
        uint32_t pw = w >> i, ph = h >> i;
        if (pw == 0 || ph == 0) return fail(TSG_ERR_INVALID_ARGUMENT, "image too small for the pyramid depth");
        HIP_TRY(small.reserve((size_t)pw * ph));
        int rc = device_resize(last, w, h, small.p, pw, ph, g, s);
        if (rc) return rc;
        rc = device_resize(small.p, pw, ph, out + (size_t)lv * npx, w, h, g, s);
        if (rc) return rc;
    }
    return TSG_OK;
}

const std::vector<int32_t>& offset_table() {
    static std::vector<int32_t> t = [] {
        struct O {
            int dx, dy;
        };
        std::vector<O> v;
        const int R = kTableRadius;
        for (int dy = -R; dy <= R; ++dy)
            for (int dx = -R; dx <= R; ++dx)
                if (dx * dx + dy * dy <= R * R) v.push_back({dx, dy});
        std::sort(v.begin(), v.end(), [](const O& a, const O& b) {
            int da = a.dx * a.dx + a.dy * a.dy, db = b.dx * b.dx + b.dy * b.dy;
            if (da != db) return da < db;
            if (a.dy != b.dy) return a.dy < b.dy;
            return a.dx < b.dx;
        });
        std::vector<int32_t> out;
        for (const O& o : v) out.push_back((int32_t)(((uint32_t)o.dx & 0xffffu) | ((uint32_t)o.dy << 16)));
        return out;
    }();
    return t;
}

inline uint32_t pack_px(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

// metric_cauchy / metric_l2 (ms.rs:1109-1120) for |a-b| = d; the reference's
// 256x256 tables depend only on |a-b| (exact in f32: the difference of two
// u8 values is exact and its square removes the sign).
void stage_luts(float cauchy, bool guided, float alpha_s, float* out512) {
    float sig2 = cauchy * cauchy;
    for (int d = 0; d < 256; ++d) {
        float x = ((float)d - 0.0f) / 255.0f;
        float x2 = x * x;
        float c = std::log1p(x2 / sig2);
        float l2 = x * x;
        out512[d] = guided ? (1.0f - alpha_s) * c : c;
        out512[256 + d] = guided ? alpha_s * l2 : 0.0f;
    }
}

}  // namespace

// ------------------------------------------------------------------------
// Host-to-device uploads of the run loop go through pinned chunks: a copy
// from pageable memory would make the host wait for the stream to drain.
// A chunk is reused once the event recorded after its last copy completes.
struct Staging {
    static constexpr int kMaxChunks = 8;
    size_t kChunk = 4u << 20;
    int kChunks = 4;
    char* buf[kMaxChunks] = {};
    hipEvent_t ev[kMaxChunks] = {};
    bool pending[kMaxChunks] = {};
    int cur = 0;
    size_t used = 0;
    hipError_t init(size_t chunk, int chunks) {
        kChunk = chunk;
        kChunks = std::min(chunks, kMaxChunks);
        for (int q = 0; q < kChunks; ++q) {
            hipError_t e = hipHostMalloc((void**)&buf[q], kChunk);
            if (e != hipSuccess) return e;
            if ((e = hipEventCreateWithFlags(&ev[q], hipEventDisableTiming)) != hipSuccess) return e;
        }
        return hipSuccess;
    }
    ~Staging() {
        for (int q = 0; q < kMaxChunks; ++q) {
            if (buf[q]) (void)hipHostFree(buf[q]);
            if (ev[q]) (void)hipEventDestroy(ev[q]);
        }
    }
    hipError_t upload(void* dst, const void* src, size_t n, hipStream_t s) {
        const char* from = static_cast<const char*>(src);
        char* to = static_cast<char*>(dst);
        while (n > 0) {
            if (used >= kChunk) {
                hipError_t e = next(s);
                if (e != hipSuccess) return e;
            }
            const size_t piece = std::min(n, kChunk - used);
            memcpy(buf[cur] + used, from, piece);
            hipError_t e = hipMemcpyAsync(to, buf[cur] + used, piece, hipMemcpyHostToDevice, s);
            if (e != hipSuccess) return e;
            used += (piece + 255) & ~(size_t)255;
            from += piece;
            to += piece;
            n -= piece;
        }
        return hipSuccess;
    }
    hipError_t next(hipStream_t s) {
        hipError_t e = hipEventRecord(ev[cur], s);
        if (e != hipSuccess) return e;
        pending[cur] = true;
        cur = (cur + 1) % kChunks;
        used = 0;
        if (pending[cur]) {
            if ((e = hipEventSynchronize(ev[cur])) != hipSuccess) return e;
            pending[cur] = false;
        }
        return hipSuccess;
    }
};

struct tsg_job;
namespace {
bool picks_step(tsg_job* j);
void picks_requeued(tsg_job* j);
}  // namespace

// Host threads that compute jobs' pick orders (the reference's swap_remove
// chain over the unresolved list) while the device works on earlier stages.
// A worker computes one stage of a job, then puts the job at the back of the
// queue: with many jobs queued, every job's first stage is ready early.
struct PicksPool {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<tsg_job*> queue;
    std::vector<std::thread> workers;
    bool stop = false;
    void submit(tsg_job* j) {
        {
            std::lock_guard<std::mutex> lk(mu);
            queue.push_back(j);
            if (workers.size() < std::min(8u, std::max(1u, std::thread::hardware_concurrency() / 2)) &&
                workers.size() < queue.size() + 1)
                workers.emplace_back([this] { loop(); });
        }
        cv.notify_one();
    }
    // true when j was still queued (and is now removed)
    bool remove(tsg_job* j) {
        std::lock_guard<std::mutex> lk(mu);
        for (auto it = queue.begin(); it != queue.end(); ++it)
            if (*it == j) {
                queue.erase(it);
                return true;
            }
        return false;
    }
    void loop() {
        for (;;) {
            tsg_job* j;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !queue.empty(); });
                if (stop) return;
                j = queue.front();
                queue.pop_front();
            }
            if (picks_step(j)) {
                std::lock_guard<std::mutex> lk(mu);
                picks_requeued(j);
                queue.push_back(j);
                cv.notify_one();
            }
        }
    }
    ~PicksPool() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        for (auto& t : workers) t.join();
    }
};

// Pinned landing area of downloads: the device writes it at full link
// speed, then host threads copy into the caller's (pageable) buffers, which
// spreads the first-touch page faults of fresh buffers over several cores.
struct Landing {
    char* buf = nullptr;
    size_t n = 0;
    hipError_t reserve(size_t bytes) {
        if (bytes <= n) return hipSuccess;
        if (buf) (void)hipHostFree(buf);
        buf = nullptr;
        n = 0;
        hipError_t e = hipHostMalloc((void**)&buf, bytes);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    ~Landing() {
        if (buf) (void)hipHostFree(buf);
    }
};

// memcpy of n bytes on up to `threads` host threads.
static void parallel_copy(void* dst, const void* src, size_t n, unsigned threads) {
    const size_t piece = 1u << 21;
    if (n <= 2 * piece || threads <= 1) {
        memcpy(dst, src, n);
        return;
    }
    const unsigned nt = (unsigned)std::min<size_t>(threads, (n + piece - 1) / piece);
    std::vector<std::thread> th;
    const size_t per = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const size_t lo = t * per, hi = std::min(n, lo + per);
        if (lo >= hi) break;
        th.emplace_back([=] { memcpy(static_cast<char*>(dst) + lo, static_cast<const char*>(src) + lo, hi - lo); });
    }
    for (auto& x : th) x.join();
}

struct tsg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Landing landing;
    hipStream_t side = nullptr;  // work that overlaps the stage loop (random candidates ahead)
    Staging staging, staging_side;  // one per stream: a chunk's reuse waits on its stream's event
    PicksPool picks;
    DBuf<int32_t> offtab;
    uint32_t offtab_n = 0;
    DBuf<float> dummy_lut;
    unsigned char* scratch = nullptr;  // this device's copy of the file-scope scratch array
    std::mutex mu;
    // launch structures stored by the autotuner, by job shape (see tuning_key)
    std::map<std::vector<uint64_t>, tsg_tuning> tunings;
};

// Threads drawing one job's pick-chain indices in a single-job run.
constexpr unsigned kPickThreads = 12;

struct tsg_job {
    tsg_ctx* ctx = nullptr;
    tsg_params prm{};
    tsg_schedule sched{};
    tsg_tuning tune{};
    uint32_t W = 0, H = 0, tiles_x = 0, tiles_y = 0;
    size_t N = 0;
    uint32_t n_ex = 0, levels = 1;
    std::vector<uint32_t> filt;
    std::vector<int> kind;
    bool guided = false;
    std::vector<std::pair<uint32_t, uint32_t>> ex_dims, g_dims;
    std::vector<std::unique_ptr<DBuf<uint32_t>>> ex, exg;  // [example] levels x w x h
    std::vector<std::unique_ptr<DBuf<uint8_t>>> smask;    // [example] (null = All)
    DBuf<uint32_t> tg;                                    // levels x W x H
    DBuf<MapDesc> maps;                                   // [level][filtered map]
    std::vector<std::vector<uint32_t>> h_finest;          // finest example pixels
    // initial state inputs
    bool inpaint = false;
    std::vector<uint8_t> inpaint_mask_r;
    uint32_t inpaint_index = 0;
    // device state
    DBuf<uint4> state;
    DBuf<uint64_t> mask;
    DBuf<uint64_t> gmask;  // tiling mode: pixels with ghost copies (ms.rs:296-331)
    DBuf<float> score;
    DBuf<uint32_t> rlist;
    // scratch
    DBuf<uint4> result;
    DBuf<uint2> rc;
    // random candidates of whole stages, generated a stage ahead on the side
    // stream (single-job runs): buffer t % 2 holds stage t, in pieces
    // published by events; done[b] marks the end of the stage that read b
    // Single-job runs compute everything content-independent a stage ahead
    // on the side stream, into buffer set t % 2: stage t's random
    // candidates, its picks in the resolved list and their ranks, the mask
    // at its end, and every item's neighbour list (the resolved set at any
    // moment follows from the pick order alone). `ready` publishes the set;
    // `done` marks the end of the stage that read it.
    struct Ahead {
        DBuf<uint2> rc;
        DBuf<int32_t> nbrs;
        DBuf<uint32_t> cnt, n_at;
        DBuf<uint64_t> mend, mnew;  // resolved mask at the stage end; the stage's picks
        DBuf<uint32_t> tmin;        // per tile: least position of its picks
        DBuf<uint64_t> dep;         // per prefetched item: neighbours that are the stage's picks
        hipEvent_t ready = nullptr, done = nullptr;
        bool done_recorded = false;
        int stage = -1;
        uint64_t n_open = 0, S_eff = 0;
        uint64_t w_end = 0;    // end of the leading run of latency-bound batches
        uint64_t pre_end = 0;  // items [0, pre_end) have neighbour lists
    } ahead[8];
    bool ahead_on = false;
    // one set per stage, every item with its neighbour list (when the stages
    // fit the budget); else two sets used in turn, lists for the leading run
    bool ahead_deep = false;
    Ahead& aset(int t) { return ahead[ahead_deep ? t : t % 2]; }
    bool in_engine = false;   // run by a tsg_batch (no single-job run structures)
    tsg_batch* owner = nullptr;  // the batch holding the job, from submission until it is taken back
    uint64_t ahead_open = 0;  // resolved count at the open of the next stage to prepare
    int ahead_next = 0;       // next stage to prepare
    bool ahead_wanted = false;  // ahead_next is due (prepared once its picks are ready)
    DBuf<float> item_score;
    DBuf<float> lut;
    DBuf<uint32_t> overflow;
    DBuf<double> st_sum;
    DBuf<unsigned long long> st_cnt;
    DBuf<double> st_part;                // per kStatsChunk items of a stage: partial sums and counts
    DBuf<unsigned long long> st_part_n;
    // under the work-queue engine: what was planned since the last launch
    struct Eng {
        std::vector<QPhase> prog;        // phases (arg: index into args)
        std::vector<BatchArgs> args;     // the next launch's stage argument blocks
        BatchArgs stage_a{};             // the open stage's block, as planned at its opening
        std::vector<uint32_t> n_at;      // prefetch windows' resolved counts (pre_base: offset)
        uint32_t stage_arg = 0;          // the open stage's entry in args
        uint32_t sharing = 1;            // jobs sharing the device when planned
        uint64_t weight = 0;             // work of the planned phases (XCD placement)
        uint64_t items = 0;              // resolve items planned for the next launch
        uint64_t quantum = 0;            // at most this many per launch (0: no bound)
        bool planned_all = false;        // every phase of the job is planned
        DBuf<float> luts;                // per stage: 512-entry cost tables
    } eng;
    DBuf<uint32_t> done_counter;
    DBuf<uint32_t> small_tab;  // batch bounds and resolved counts of small runs
    // neighbour lists computed ahead of latency-bound batches
    DBuf<uint32_t> rank;       // per pixel: position in the resolved list (UINT32_MAX: none)
    DBuf<uint64_t> mend;       // pixels resolved inside a prefetch window
    DBuf<uint32_t> tmin;       // per tile: least resolved-list position among them
    DBuf<uint32_t> pre_n;      // per window item: resolved count at its batch start
    DBuf<int32_t> pre_nbrs;    // per window item: k packed offsets
    DBuf<uint32_t> pre_cnt;
    // dependency-ordered runs: per pixel the epoch of the run that committed
    // it, the runs' ticket counter and its value as the next run starts
    DBuf<uint32_t> flow_flag;
    DBuf<uint32_t> flow_ticket;
    DBuf<uint4> state_old;  // redo runs: the state as the redo segment opened
    // deferred commits: records of the last batches, their batch serials, and
    // three pending masks used in turn (serial % 3)
    DBuf<uint4> state2;
    DBuf<uint32_t> stamp;
    DBuf<uint64_t> pend;
    uint32_t serial = 2;
    uint32_t flow_epoch = 0, flow_tickets = 0;
    size_t small_used = 0;
    // host bookkeeping
    std::vector<uint32_t> unresolved;
    size_t n_res = 0, locked = 0, N0 = 0;
    std::vector<uint64_t> st_items;
    std::vector<std::vector<uint32_t>> picks_keep;
    std::vector<std::vector<float>> lut_keep;
    bool ran = false;
    bool state_ready = false;
    size_t rc_budget = 32u << 20;  // bytes of random candidates generated per chunk
    // pick orders computed by the context's pool: picks_ready[t] and
    // picks_keep[t] are published under pk_mu
    std::vector<uint8_t> picks_ready;
    std::mutex pk_mu;
    std::condition_variable pk_cv;
    int pk_state = 0;  // 0 not submitted, 1 queued, 2 computing, 3 finished
    bool pk_cancel = false;
    unsigned pk_threads = kPickThreads;  // threads of one pick chain's index draws (capped by the host's cores)
    uint64_t pk_start_res = 0;
    uint64_t overall_total = 0, overall_cur = 0;
    hipEvent_t ev_run0 = nullptr, ev_run1 = nullptr;
    // cursor of the resumable stage driver
    struct Step {
        bool active = false, open = false;
        int t = 0;
        uint32_t level = 0;
        uint64_t pss = 0, S = 0, S_eff = 0, redo = 0, redo_eff = 0, i = 0;
        uint64_t rc_lo = 0, rc_hi = 0;
        BatchArgs a{};
    } step;
    double run_ms = 0.0;
    uint64_t n_batches = 0;
    // optional per-kernel device timing (events on the job's stream)
    bool profile = false;
    int profile_level = 0;
    DBuf<unsigned long long> clk;  // per-item phase clocks of the first stage (level 2)
    struct Timed {
        int kind;
        uint64_t items;
        hipEvent_t a, b;
        int64_t slot;  // device-clock span slot (resolve kernels), -1: none
    };
    DBuf<unsigned long long> launch_clk;  // {start, end} pairs of resolve launches (profiling)
    size_t launch_slots = 0;
    std::vector<Timed> timed;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_next = 0;
    double kern_ms[7] = {};
    uint64_t kern_items[7] = {};
    uint64_t kern_launches[7] = {};
    ~tsg_job() {
        // the pick worker must be done with this job
        if (ctx && !ctx->picks.remove(this)) {
            std::unique_lock<std::mutex> lk(pk_mu);
            pk_cancel = true;
            pk_cv.wait(lk, [&] { return pk_state != 1 && pk_state != 2; });
        }
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        for (Ahead& a : ahead) {
            if (a.ready) (void)hipEventDestroy(a.ready);
            if (a.done) (void)hipEventDestroy(a.done);
        }
        if (ev_run0) (void)hipEventDestroy(ev_run0);
        if (ev_run1) (void)hipEventDestroy(ev_run1);
    }
};

namespace {

uint64_t stage_size(const tsg_job* j, int s) {
    float v = std::pow(j->prm.backtrack_percent, (float)s) * (float)j->N0;
    if (!(v > 0.0f)) return 0;
    if (v >= 18446744073709551615.0f) return UINT64_MAX;
    return (uint64_t)v;
}

// pick_random_unresolved for items [i0, i0 + n): index draws in parallel
// (each depends only on its seed and the list length at that step), then
// the swap_remove chain in order.
std::vector<uint32_t> pick_new(tsg_job* j, uint64_t stage_seed, uint64_t i0, uint64_t n) {
    size_t L0 = j->unresolved.size();
    n = std::min<uint64_t>(n, L0);
    std::vector<uint64_t> idx(n);
    // leave the job's own thread a core
    const unsigned hw = std::thread::hardware_concurrency();
    unsigned nt = std::max(1u, std::min(j->pk_threads, hw > 2 ? hw - 1 : 1u));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t] {
            for (uint64_t q = t; q < n; q += nt) {
                Pcg32 g = Pcg32::seed_from_u64(stage_seed + i0 + q);
                idx[q] = range_u64(g, 0, L0 - q);
            }
        });
    }
    for (auto& x : th) x.join();
    // the swap_remove chain; every index is known, so the line each step
    // touches is fetched a few dozen steps early
    std::vector<uint32_t> out(n);
    uint32_t* u = j->unresolved.data();
    size_t len = L0;
    for (uint64_t q = 0; q < n; ++q) {
        if (q + 24 < n) __builtin_prefetch(u + idx[q + 24], 1, 3);
        const uint64_t k = idx[q];
        out[q] = u[k];
        u[k] = u[len - 1];
        --len;
    }
    j->unresolved.resize(len);
    return out;
}

// resolve_at_random source (ms.rs:447-475): three identically seeded streams.
void random_source(uint64_t seed, const std::vector<std::pair<uint32_t, uint32_t>>& dims, uint32_t& m, uint32_t& x,
                   uint32_t& y) {
    Pcg32 a = Pcg32::seed_from_u64(seed);
    m = (uint32_t)range_u64(a, 0, dims.size());
    Pcg32 b = Pcg32::seed_from_u64(seed);
    x = range_u32(b, 0, dims[m].first);
    Pcg32 c = Pcg32::seed_from_u64(seed);
    y = range_u32(c, 0, dims[m].second);
}

int check_range(float v, float lo, float hi, const char* name) {
    if (v < lo || v > hi) {
        return fail(TSG_ERR_INVALID_RANGE, std::string("parameter '") + name + "' - value '" + std::to_string(v) +
                                               "' is outside the range of " + std::to_string(lo) + "-" +
                                               std::to_string(hi));
    }
    return TSG_OK;
}

// Initial state: Generator::new / new_from_inpaint (ms.rs:220-293) and the
// random init of Session::run (session.rs:42-52).
int init_state(tsg_job* j) {
    const size_t N = j->N;
    std::vector<uint4> st(N, make_uint4(0, 0, 0, 0));
    std::vector<uint64_t> mask((size_t)j->tiles_x * j->tiles_y, 0);
    std::vector<uint32_t> rl;
    rl.reserve(N);
    auto set_bit = [&](uint32_t f) {
        uint32_t x = f % j->W, y = f / j->W;
        mask[(y >> 3) * j->tiles_x + (x >> 3)] |= 1ull << (((y & 7u) << 3) | (x & 7u));
    };
    j->unresolved.clear();
    if (j->inpaint) {
        const std::vector<uint32_t>& col = j->h_finest[j->inpaint_index];
        for (size_t i = 0; i < N; ++i) {
            st[i].x = col[i];
            if (j->inpaint_mask_r[i] < 255) {
                j->unresolved.push_back((uint32_t)i);
            } else {
                uint32_t x = (uint32_t)(i % j->W), y = (uint32_t)(i / j->W);
                st[i] = make_uint4(col[i], x | (y << 16), j->inpaint_index, 0u);
                set_bit((uint32_t)i);
                rl.push_back((uint32_t)i);
            }
        }
    } else {
        j->unresolved.resize(N);
        for (size_t i = 0; i < N; ++i) j->unresolved[i] = (uint32_t)i;
    }
    j->locked = rl.size();
    if (j->prm.random_init != UINT64_MAX) {
        for (uint64_t i = 0; i < j->prm.random_init; ++i) {
            if (j->unresolved.empty()) continue;
            Pcg32 g = Pcg32::seed_from_u64(j->prm.seed + i);
            uint64_t k = range_u64(g, 0, j->unresolved.size());
            uint32_t f = j->unresolved[k];
            j->unresolved[k] = j->unresolved.back();
            j->unresolved.pop_back();
            uint32_t m, x, y;
            random_source(j->prm.seed + i + (uint64_t)f, j->ex_dims, m, x, y);
            uint32_t c = j->h_finest[m][(size_t)y * j->ex_dims[m].first + x];
            st[f] = make_uint4(c, x | (y << 16), m | (m << 16), f);
            set_bit(f);
            rl.push_back(f);
        }
        j->locked += j->prm.random_init;
        if (j->locked > rl.size()) return fail(TSG_ERR_INVALID_ARGUMENT, "random_init exceeds the pixels to resolve");
    }
    j->n_res = rl.size();
    j->N0 = j->unresolved.size();
    hipStream_t s = j->ctx->stream;
    HIP_TRY(hipMemcpyAsync(j->state.p, st.data(), N * sizeof(uint4), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(j->mask.p, mask.data(), mask.size() * 8, hipMemcpyHostToDevice, s));
    // every pixel resolved before the stages starts with its ghost copies (ms.rs:296-331)
    if (j->prm.tiling_mode) HIP_TRY(hipMemcpyAsync(j->gmask.p, mask.data(), mask.size() * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(j->score.p, 0, N * sizeof(float), s));
    if (!rl.empty()) HIP_TRY(hipMemcpyAsync(j->rlist.p, rl.data(), rl.size() * 4, hipMemcpyHostToDevice, s));
    std::vector<uint32_t> rank(N, UINT32_MAX);
    for (size_t q = 0; q < rl.size(); ++q) rank[rl[q]] = (uint32_t)q;
    HIP_TRY(hipMemcpyAsync(j->rank.p, rank.data(), N * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return TSG_OK;
}

// End of the batch starting at item i of stage t when n_res pixels are resolved.
uint64_t batch_end_at(const tsg_job* j, int t, uint64_t i, uint64_t n_res, uint64_t redo, uint64_t S_eff) {
    uint64_t redo_eff = std::min<uint64_t>(redo, S_eff);
    uint64_t seg_end = i < redo_eff ? redo_eff : S_eff;
    uint64_t b;
    if (j->sched.px_per_batch > 0) {
        b = j->sched.px_per_batch;
    } else {
        uint64_t base = n_res > j->locked ? n_res - j->locked : 0;
        const uint32_t ns = (t == 0 && j->sched.first_new_shift) ? j->sched.first_new_shift : j->sched.new_shift;
        b = base >> (i < redo_eff ? j->sched.redo_shift : ns);
        if (b < 1) b = 1;
        if (j->sched.max_batch > 0 && b > j->sched.max_batch) b = j->sched.max_batch;
    }
    return std::min(i + b, seg_end);
}
uint64_t batch_end(const tsg_job* j, uint64_t i, uint64_t redo, uint64_t S_eff) {
    return batch_end_at(j, j->step.t, i, j->n_res, redo, S_eff);
}

enum { K_RESOLVE = 0, K_COMMIT = 1, K_RAND = 2, K_RECOLOR = 3, K_STATS = 4, K_PREFETCH = 5, K_RUN = 6 };
constexpr int kKinds = 7;
constexpr size_t kEngineRcBudget = 4u << 20;  // random-candidate chunk of a job under the engine
constexpr uint64_t kPrefetchCap = 65536;       // items per prefetched kNN window
constexpr uint32_t kFlowSpin = 1u << 18;       // polls of one dependency wait before a run gives up

hipError_t next_event(tsg_job* j, hipEvent_t* e) {
    if (j->ev_next == j->ev_pool.size()) {
        hipEvent_t n;
        hipError_t r = hipEventCreate(&n);
        if (r != hipSuccess) return r;
        j->ev_pool.push_back(n);
    }
    *e = j->ev_pool[j->ev_next++];
    return hipSuccess;
}

// Runs `launch` (on stream s), bracketed by events when profiling.
template <class F>
hipError_t timed_launch_on(tsg_job* j, hipStream_t s, int kind, uint64_t items, F&& launch) {
    if (!j->profile) return launch();
    tsg_job::Timed t{kind, items, nullptr, nullptr, -1};
    hipError_t e;
    if (j->profile_level < 0) {
        // spans only: the resolve kernels' device-clock spans, no events
        if ((kind != K_RESOLVE && kind != K_RUN) || j->launch_slots >= j->launch_clk.n / 2) return launch();
        t.slot = (int64_t)j->launch_slots++;
        j->step.a.dbg_launch = j->launch_clk.p + 2 * t.slot;
        e = launch();
        j->step.a.dbg_launch = nullptr;
        if (e != hipSuccess) return e;
        j->timed.push_back(t);
        return hipSuccess;
    }
    if ((e = next_event(j, &t.a)) != hipSuccess) return e;
    if ((e = next_event(j, &t.b)) != hipSuccess) return e;
    if ((kind == K_RESOLVE || kind == K_RUN) && j->launch_slots < j->launch_clk.n / 2) {
        // the kernel's workgroups record its execution span on the device clock
        t.slot = (int64_t)j->launch_slots++;
        j->step.a.dbg_launch = j->launch_clk.p + 2 * t.slot;
        set_launch_events(t.a, t.b);
        e = launch();
        set_launch_events(nullptr, nullptr);
        j->step.a.dbg_launch = nullptr;
        if (e != hipSuccess) return e;
    } else {
        if ((e = hipEventRecord(t.a, s)) != hipSuccess) return e;
        if ((e = launch()) != hipSuccess) return e;
        if ((e = hipEventRecord(t.b, s)) != hipSuccess) return e;
    }
    j->timed.push_back(t);
    return hipSuccess;
}
template <class F>
hipError_t timed_launch(tsg_job* j, int kind, uint64_t items, F&& launch) {
    return timed_launch_on(j, j->ctx->stream, kind, items, launch);
}

constexpr uint64_t kAheadBudget = 16ull << 30;  // device bytes of the stage-ahead buffers

// ---- resumable stage driver (Generator::resolve, ms.rs:786-1050) ----

// Pick order of the next stage's new pixels (the chain depends only on
// counts and seeds, not on image content), published as soon as it is done.
// Runs on a pool thread; true when stages remain (the pool queues the job
// again), false when every stage is done or the job was cancelled.
bool picks_step(tsg_job* j) {
    const int P = (int)j->prm.backtrack_stages;
    int t = 0;
    uint64_t planned = j->pk_start_res;
    {
        std::lock_guard<std::mutex> lk(j->pk_mu);
        if (j->pk_cancel) {
            j->pk_state = 3;
            j->pk_cv.notify_all();
            return false;
        }
        j->pk_state = 2;
        while (t <= P && j->picks_ready[t]) planned += j->picks_keep[t++].size();
    }
    if (t <= P) {
        const int ps = P - t;
        const uint64_t pss = (uint64_t)Pcg32::seed_from_u64(j->prm.seed + (uint64_t)ps).next_u32();
        const uint64_t S = stage_size(j, ps);
        const uint64_t redo = planned - j->locked;
        std::vector<uint32_t> picks;
        if (S > redo) picks = pick_new(j, pss, redo, S - redo);
        std::lock_guard<std::mutex> lk(j->pk_mu);
        j->picks_keep[t] = std::move(picks);
        j->picks_ready[t] = 1;
        ++t;
    }
    std::lock_guard<std::mutex> lk(j->pk_mu);
    const bool more = t <= P && !j->pk_cancel;
    if (!more) j->pk_state = 3;
    j->pk_cv.notify_all();
    return more;
}

// The job goes back into the pool's queue for its next stage.
void picks_requeued(tsg_job* j) {
    std::lock_guard<std::mutex> lk(j->pk_mu);
    j->pk_state = 1;
}

bool picks_ready_now(tsg_job* j, int t) {
    std::lock_guard<std::mutex> lk(j->pk_mu);
    return j->picks_ready[t] != 0;
}

int wait_picks(tsg_job* j, int t) {
    std::unique_lock<std::mutex> lk(j->pk_mu);
    j->pk_cv.wait(lk, [&] { return j->picks_ready[t] || j->pk_state == 3; });
    if (!j->picks_ready[t]) return fail(TSG_ERR_DEVICE, "pick order computation stopped");
    return TSG_OK;
}

// Device bytes of the stage-ahead buffers: two sets sized for the largest
// stage, or (deep) one set per stage.
uint64_t ahead_set_bytes(const tsg_job* j, uint64_t S) {
    return std::max<uint64_t>(1, S) * (j->prm.random_sample_locations * sizeof(uint2) + j->prm.nearest_neighbors * 4 + 16) +
           (uint64_t)j->tiles_x * j->tiles_y * 20;
}
uint64_t ahead_bytes(const tsg_job* j) { return 2 * ahead_set_bytes(j, stage_size(j, 0)); }
uint64_t ahead_bytes_deep(const tsg_job* j) {
    uint64_t b = 0;
    for (int s = 0; s <= (int)j->prm.backtrack_stages; ++s) b += ahead_set_bytes(j, stage_size(j, s));
    return b;
}

// Prepares stage ahead_next on the side stream (see tsg_job::Ahead). The
// counts follow the stage driver: a stage opens with `ahead_open` resolved
// pixels, redoes those past `locked`, and appends its picks.
int ahead_stage(tsg_job* j) {
    const int P = (int)j->prm.backtrack_stages;
    const int t = j->ahead_next;
    if (t > P) return TSG_OK;
    int rc = wait_picks(j, t);
    if (rc) return rc;
    tsg_job::Ahead& A = j->aset(t);
    hipStream_t side = j->ctx->side;
    const uint32_t nf = (uint32_t)j->filt.size();
    const uint32_t m = (uint32_t)j->prm.random_sample_locations;
    const uint32_t k = j->prm.nearest_neighbors;
    const int ps = P - t;
    const uint32_t level = t == 0 ? 0u : (uint32_t)std::min(t, P - 1);
    const uint64_t S = stage_size(j, ps);
    const uint64_t pss = (uint64_t)Pcg32::seed_from_u64(j->prm.seed + (uint64_t)ps).next_u32();
    const uint64_t n_open = j->ahead_open;
    const uint64_t redo = n_open - j->locked;
    const std::vector<uint32_t>& picks = j->picks_keep[t];
    const uint64_t S_eff = S > redo ? redo + picks.size() : S;
    const uint64_t redo_eff = std::min(redo, S_eff);
    if (A.done_recorded) HIP_TRY(hipStreamWaitEvent(side, A.done, 0));
    const size_t tiles = (size_t)j->tiles_x * j->tiles_y;
    if (j->ahead_deep) {
        // this stage's own set
        const uint64_t n = std::max<uint64_t>(1, S_eff);
        HIP_TRY(A.rc.reserve(n * m));
        HIP_TRY(A.nbrs.reserve(n * k));
        HIP_TRY(A.cnt.reserve(n));
        HIP_TRY(A.dep.reserve(n));
        HIP_TRY(A.mend.reserve(tiles));
        HIP_TRY(A.mnew.reserve(tiles));
        HIP_TRY(A.tmin.reserve(tiles));
        if (!A.ready) HIP_TRY(hipEventCreateWithFlags(&A.ready, hipEventDisableTiming));
    }
    // random candidates of every item (ms.rs:549-599: stage seed, item index, level maps)
    if (S_eff) {
        RandArgs ra{j->maps.p + (size_t)level * nf, nf, pss, 0u, (uint32_t)S_eff, m, A.rc.p, j->overflow.p};
        HIP_TRY(timed_launch_on(j, side, K_RAND, S_eff, [&] { return launch_rand(ra, side); }));
    }
    // picks into the resolved list, their ranks, the stage's picks as a mask
    // (with each tile's least rank) and the mask at the stage end
    const uint64_t* open_mask = t == 0 ? j->mask.p : j->aset(t - 1).mend.p;
    HIP_TRY(hipMemcpyAsync(A.mend.p, open_mask, tiles * 8, hipMemcpyDeviceToDevice, side));
    HIP_TRY(hipMemsetAsync(A.mnew.p, 0, tiles * 8, side));
    HIP_TRY(hipMemsetAsync(A.tmin.p, 0xff, tiles * 4, side));
    if (!picks.empty()) {
        const uint32_t lo = (uint32_t)n_open, hi = (uint32_t)(n_open + picks.size());
        HIP_TRY(j->ctx->staging_side.upload(j->rlist.p + n_open, picks.data(), picks.size() * 4, side));
        HIP_TRY(launch_rank(j->rlist.p, lo, hi, j->rank.p, side));
        HIP_TRY(launch_mask_or(j->rlist.p, lo, hi, j->W, j->tiles_x, A.mnew.p, A.tmin.p, side));
        HIP_TRY(launch_mask_or(j->rlist.p, lo, hi, j->W, j->tiles_x, A.mend.p, nullptr, side));
    }
    // the stage's batches (start item, resolved count at the start); the
    // neighbour lists of the items in the leading latency-bound batches
    // (deep: of every item; the rest search in-kernel)
    std::vector<uint32_t> tab;  // bounds (nb + 1), then resolved counts (nb)
    std::vector<uint32_t> nres_b;
    uint64_t i = 0, nres = n_open, w_end = 0;
    bool window = true;
    while (i < S_eff) {
        const uint64_t e = batch_end_at(j, t, i, nres, redo, S_eff);
        tab.push_back((uint32_t)i);
        nres_b.push_back((uint32_t)nres);
        if (window && e - i <= j->tune.prefetch_max) w_end = e;
        else window = false;
        if (e > redo_eff) nres += e - std::max(i, redo_eff);
        i = e;
    }
    const uint32_t nb = (uint32_t)nres_b.size();
    tab.push_back((uint32_t)S_eff);
    tab.insert(tab.end(), nres_b.begin(), nres_b.end());
    const uint64_t pre_end = j->ahead_deep ? S_eff : w_end;
    if (pre_end) {
        HIP_TRY(A.n_at.reserve(tab.size()));
        HIP_TRY(j->ctx->staging_side.upload(A.n_at.p, tab.data(), tab.size() * 4, side));
        PrefetchArgs pa{};
        pa.W = j->W;
        pa.H = j->H;
        pa.tiles_x = j->tiles_x;
        pa.tiles_y = j->tiles_y;
        pa.mask = open_mask;
        pa.mask_new = A.mnew.p;
        pa.tile_min = A.tmin.p;
        pa.rank = j->rank.p;
        pa.rlist = j->rlist.p;
        pa.work = j->rlist.p + j->locked;
        pa.bounds = A.n_at.p;
        pa.bnres = A.n_at.p + nb + 1;
        pa.nb = nb;
        pa.redo_end = (uint32_t)redo_eff;
        pa.locked = (uint32_t)j->locked;
        pa.offtab = j->ctx->offtab.p;
        pa.offtab_n = j->ctx->offtab_n;
        pa.k = k;
        pa.item0 = 0;
        pa.item1 = (uint32_t)pre_end;
        pa.out_nbrs = A.nbrs.p;
        pa.out_cnt = A.cnt.p;
        pa.out_dep = A.dep.p;
        HIP_TRY(timed_launch_on(j, side, K_PREFETCH, pre_end, [&] { return launch_knn_prefetch(pa, side); }));
    }
    HIP_TRY(hipEventRecord(A.ready, side));
    A.stage = t;
    A.n_open = n_open;
    A.S_eff = S_eff;
    A.w_end = w_end;
    A.pre_end = pre_end;
    j->ahead_open = nres;
    j->ahead_next = t + 1;
    return TSG_OK;
}

bool picks_available(tsg_job* j, int t) {
    std::lock_guard<std::mutex> lk(j->pk_mu);
    return j->picks_ready[t] || j->pk_state == 3;
}

// Prepares the due stage when its picks are ready, or (block) waits for them.
// The stage driver polls between launches, so the host keeps feeding the
// device while the pick threads finish.
int ahead_poll(tsg_job* j, bool block) {
    if (!j->ahead_wanted) return TSG_OK;
    if (j->ahead_next > (int)j->prm.backtrack_stages) {
        j->ahead_wanted = false;
        return TSG_OK;
    }
    if (!block && !picks_available(j, j->ahead_next)) return TSG_OK;
    j->ahead_wanted = false;
    return ahead_stage(j);
}

int step_begin(tsg_job* j) {
    hipStream_t s = j->ctx->stream;
    if (!j->state_ready) return fail(TSG_ERR_INVALID_ARGUMENT, "job state was not initialised");
    if (j->ran || j->step.active) return fail(TSG_ERR_INVALID_ARGUMENT, "a job runs once (Session::run consumes the session)");
    j->timed.clear();
    j->ev_next = 0;
    j->launch_slots = 0;
    if (j->profile) {
        // spans start empty: {~start, end} = {0, 0}; 8192 timed launches
        HIP_TRY(j->launch_clk.reserve(2 * 8192));
        HIP_TRY(hipMemsetAsync(j->launch_clk.p, 0, j->launch_clk.n * 8, s));
    }
    for (int q = 0; q < kKinds; ++q) j->kern_ms[q] = 0.0, j->kern_items[q] = 0, j->kern_launches[q] = 0;
    const int P = (int)j->prm.backtrack_stages;
    j->st_items.assign(P + 1, 0);
    HIP_TRY(hipMemsetAsync(j->st_sum.p, 0, (P + 1) * sizeof(double), s));
    HIP_TRY(hipMemsetAsync(j->st_cnt.p, 0, (P + 1) * sizeof(unsigned long long), s));
    HIP_TRY(hipMemsetAsync(j->overflow.p, 0, 4, s));
    HIP_TRY(hipMemsetAsync(j->done_counter.p, 0, 4, s));
    {
        std::lock_guard<std::mutex> lk(j->pk_mu);
        j->picks_keep.assign(P + 1, {});
        j->picks_ready.assign(P + 1, 0);
        j->pk_start_res = j->n_res;
        j->pk_state = 1;
    }
    j->ctx->picks.submit(j);
    j->lut_keep.assign(P + 1, std::vector<float>(512));
    j->small_used = 0;
    j->n_batches = 0;
    // Table arena of the small-batch runs: a run of nb batches takes 2 nb + 1
    // entries. Batches of at most 16 items occur only while
    // (resolved - locked) >> shift <= 16, i.e. for at most 17 << shift
    // batches per segment, or for every batch when px_per_batch <= 16.
    uint64_t items = 0;
    for (int t = 0; t <= P; ++t) items += stage_size(j, P - t);
    uint64_t bound = items;
    const uint64_t run_max = std::max<uint64_t>(j->tune.small_max, j->tune.coop_wgs ? j->tune.coop_max : 0);
    if (j->sched.px_per_batch == 0 || j->sched.px_per_batch > run_max) {
        uint32_t sh = std::max(std::max(j->sched.redo_shift, j->sched.new_shift), j->sched.first_new_shift);
        uint64_t per_seg = sh >= 40 ? items : (run_max + 1) << sh;
        bound = std::min<uint64_t>(items, 2ull * (P + 1) * per_seg);
    }
    if (!j->in_engine) HIP_TRY(j->small_tab.reserve(3 * bound + 4 * (P + 1) + 16));
    // largest batch: a fraction of the resolved count, a fixed size, or the cap
    uint64_t maxb = j->sched.px_per_batch ? j->sched.px_per_batch
                                          : (uint64_t)j->N >> std::min(j->sched.redo_shift, j->sched.new_shift);
    if (j->sched.max_batch) maxb = std::min<uint64_t>(maxb, j->sched.max_batch);
    maxb = std::max<uint64_t>(maxb + 1, small_batch_max());
    HIP_TRY(j->result.reserve(std::min<uint64_t>(maxb, j->N)));
    j->overall_total = items;
    j->overall_cur = 0;
    j->step = tsg_job::Step{};
    j->step.active = true;
    HIP_TRY(hipEventCreate(&j->ev_run0));
    HIP_TRY(hipEventCreate(&j->ev_run1));
    HIP_TRY(hipEventRecord(j->ev_run0, s));
    return TSG_OK;
}

// Stage preamble (ms.rs:786-858): level, recolour, seed, size, alpha, LUTs, picks.
int step_open_stage(tsg_job* j) {
    hipStream_t s = j->ctx->stream;
    tsg_job::Step& st = j->step;
    const int P = (int)j->prm.backtrack_stages;
    const int t = st.t;
    const int ps = P - t;
    const uint32_t nf = (uint32_t)j->filt.size();
    const uint32_t m = (uint32_t)j->prm.random_sample_locations;
    st.level = t == 0 ? 0u : (uint32_t)std::min(t, P - 1);
    const MapDesc* maps = j->maps.p + (size_t)st.level * nf;
    if (st.level > 0) {
        RecolorArgs ra{maps, j->rlist.p, (uint32_t)j->n_res, j->state.p};
        HIP_TRY(timed_launch(j, K_RECOLOR, j->n_res, [&] { return launch_recolor(ra, s); }));
    }
    st.pss = (uint64_t)Pcg32::seed_from_u64(j->prm.seed + (uint64_t)ps).next_u32();
    st.S = stage_size(j, ps);
    st.redo = j->n_res - j->locked;
    float alpha_s = 0.0f;
    if (j->guided && ps > 0) {
        float b = j->prm.guide_alpha * (1.0f - ((float)j->n_res / (float)j->N0));
        alpha_s = (b * b) * b;
    }
    stage_luts(j->prm.cauchy_dispersion, j->guided, alpha_s, j->lut_keep[t].data());
    uint32_t lut_nan = 0;
    for (float v : j->lut_keep[t]) lut_nan |= (uint32_t)std::isnan(v);
    HIP_TRY(j->ctx->staging.upload(j->lut.p, j->lut_keep[t].data(), 512 * 4, s));
    int rc = wait_picks(j, t);
    if (rc) return rc;
    st.S_eff = st.S;
    if (st.S > st.redo) {
        st.S_eff = st.redo + j->picks_keep[t].size();
        if (!j->picks_keep[t].empty() && !j->ahead_on) {
            HIP_TRY(j->ctx->staging.upload(j->rlist.p + j->locked + st.redo, j->picks_keep[t].data(),
                                           j->picks_keep[t].size() * 4, s));
            const uint32_t lo = (uint32_t)(j->locked + st.redo);
            HIP_TRY(launch_rank(j->rlist.p, lo, lo + (uint32_t)j->picks_keep[t].size(), j->rank.p, s));
        }
    }
    st.redo_eff = std::min(st.redo, st.S_eff);
    st.i = 0;
    st.rc_lo = st.rc_hi = 0;
    if (j->ahead_on) {
        // this stage's set was prepared during the previous stage; the next
        // stage's preparation overlaps this one
        const tsg_job::Ahead& A = j->aset(t);
        if (A.stage != t || A.n_open != j->n_res || A.S_eff != st.S_eff)
            return fail(TSG_ERR_DEVICE, "stage-ahead buffers out of step");
        HIP_TRY(hipStreamWaitEvent(s, A.ready, 0));
        j->ahead_wanted = true;
        int rc2 = ahead_poll(j, false);
        if (rc2) return rc2;
    }
    BatchArgs& a = st.a;
    a = BatchArgs{};
    a.W = j->W;
    a.H = j->H;
    a.tiles_x = j->tiles_x;
    a.tiles_y = j->tiles_y;
    a.mask = j->mask.p;
    a.state = j->state.p;
    a.tg = j->guided ? j->tg.p + (size_t)st.level * j->N : nullptr;
    a.maps = maps;
    a.n_maps = nf;
    a.work = j->rlist.p + j->locked;
    a.rlist = j->rlist.p;
    a.m = m;
    a.offtab = j->ctx->offtab.p;
    a.offtab_n = j->ctx->offtab_n;
    a.lut = j->lut.p;
    a.stage_seed = st.pss;
    a.k = j->prm.nearest_neighbors;
    a.item_score = j->item_score.p;
    a.lut_nan = lut_nan;
    a.first_new = (uint32_t)st.redo_eff;
    a.done_counter = j->done_counter.p;
    a.dbg_clk = (j->profile_level >= 2 && t == j->profile_level - 2) ? j->clk.p : nullptr;
    a.state_w = j->state.p;
    a.mask_w = j->mask.p;
    a.score_w = j->score.p;
    a.result = j->result.p;
    if (j->prm.tiling_mode) {
        // ghost margins: 5% of each side (ms.rs:296-331)
        a.gmask = j->gmask.p;
        a.xl = (int32_t)((float)j->W * 0.05f);
        a.xr = (int32_t)j->W - a.xl;
        a.yb = (int32_t)((float)j->H * 0.05f);
        a.yt = (int32_t)j->H - a.yb;
    }
    if (j->ahead_on) {
        // every item of the stage has its candidates and neighbour list
        a.rcands = j->aset(t).rc.p;
        a.rc_base = 0;
        st.rc_lo = 0;
        st.rc_hi = st.S_eff;
        a.pre_nbrs = j->aset(t).nbrs.p;
        a.pre_cnt = j->aset(t).cnt.p;
        a.pre_base = 0;
        a.pre_end = (uint32_t)j->aset(t).pre_end;
    }
    st.open = true;
    return TSG_OK;
}

// Random candidates must cover items [lo, hi) of the open stage. When they
// do not, plans a generation (into *ra) and points the batch arguments at
// it; the caller launches it before the batch.
int step_plan_rc(tsg_job* j, uint64_t lo, uint64_t hi, RandArgs* ra, bool* need) {
    tsg_job::Step& st = j->step;
    *need = false;
    if (lo >= st.rc_lo && hi <= st.rc_hi) return TSG_OK;
    const uint32_t m = (uint32_t)j->prm.random_sample_locations;
    uint64_t cap_items = std::max<uint64_t>(hi - lo, (uint64_t)j->rc_budget / (8ull * std::max(1u, m)));
    st.rc_lo = lo;
    st.rc_hi = std::min<uint64_t>(st.S_eff, lo + cap_items);
    HIP_TRY(j->rc.reserve((st.rc_hi - st.rc_lo) * m));
    *ra = RandArgs{st.a.maps, st.a.n_maps, st.pss, (uint32_t)st.rc_lo, (uint32_t)st.rc_hi, m, j->rc.p, j->overflow.p};
    st.a.rcands = j->rc.p;
    st.a.rc_base = (uint32_t)st.rc_lo;
    *need = true;
    return TSG_OK;
}

int step_ensure_rc(tsg_job* j, uint64_t lo, uint64_t hi) {
    RandArgs ra;
    bool need;
    int rc = step_plan_rc(j, lo, hi, &ra, &need);
    if (rc || !need) return rc;
    HIP_TRY(timed_launch(j, K_RAND, ra.item1 - ra.item0, [&] { return launch_rand(ra, j->ctx->stream); }));
    return TSG_OK;
}

// Next batch [i0, i1) of the open stage; false at the stage end.
bool step_peek(tsg_job* j, uint64_t& i0, uint64_t& i1) {
    tsg_job::Step& st = j->step;
    if (st.i >= st.S_eff) return false;
    i0 = st.i;
    i1 = batch_end(j, st.i, st.redo, st.S_eff);
    return true;
}

// Whether batch [i0, i1) is latency-bound and its neighbour lists are not
// computed yet.
bool step_wants_prefetch(const tsg_job* j, uint64_t i0, uint64_t i1) {
    const BatchArgs& a = j->step.a;
    // tiling mode searches in-kernel: a window's picks may or may not get ghost copies
    if (j->ahead_on || j->prm.tiling_mode) return false;
    if (i1 - i0 > j->tune.prefetch_max) return false;
    return !(a.pre_nbrs && i0 >= a.pre_base && i1 <= a.pre_end);
}

// Neighbour lists of the window of consecutive latency-bound batches that
// starts at i0 (each at most prefetch_max items, inside i0's segment, at most
// kPrefetchCap items), in one launch: item i searches the resolved set of
// its batch start, i.e. the pixels listed before position n_at(i) of the
// resolved list. The mask holds the window's pixels too; rank filters them.
int step_prefetch(tsg_job* j, uint64_t i0) {
    hipStream_t s = j->ctx->stream;
    tsg_job::Step& st = j->step;
    std::vector<uint32_t> n_at;
    const bool redo = i0 < st.redo_eff;
    const uint64_t seg_end = redo ? st.redo_eff : st.S_eff;
    uint64_t i = i0, nres = j->n_res;
    while (i < seg_end) {
        const uint64_t e = batch_end_at(j, st.t, i, nres, st.redo, st.S_eff);
        if (e - i > j->tune.prefetch_max || e - i0 > kPrefetchCap) break;
        n_at.insert(n_at.end(), e - i, (uint32_t)nres);
        if (e > st.redo_eff) nres += e - std::max(i, st.redo_eff);
        i = e;
    }
    const uint64_t w1 = i;
    if (w1 <= i0) return TSG_OK;
    const uint32_t k = j->prm.nearest_neighbors;
    const uint64_t n = w1 - i0;
    HIP_TRY(j->pre_n.reserve(std::min<uint64_t>(kPrefetchCap, j->N)));
    HIP_TRY(j->pre_cnt.reserve(std::min<uint64_t>(kPrefetchCap, j->N)));
    HIP_TRY(j->pre_nbrs.reserve(std::min<uint64_t>(kPrefetchCap, j->N) * k));
    HIP_TRY(j->ctx->staging.upload(j->pre_n.p, n_at.data(), n * 4, s));
    PrefetchArgs pa{};
    pa.W = j->W;
    pa.H = j->H;
    pa.tiles_x = j->tiles_x;
    pa.tiles_y = j->tiles_y;
    pa.rlist = j->rlist.p;
    pa.work = j->rlist.p + j->locked;
    pa.n_at = j->pre_n.p;
    pa.offtab = j->ctx->offtab.p;
    pa.offtab_n = j->ctx->offtab_n;
    pa.k = k;
    pa.item0 = (uint32_t)i0;
    pa.item1 = (uint32_t)w1;
    pa.out_nbrs = j->pre_nbrs.p;
    pa.out_cnt = j->pre_cnt.p;
    // the live mask as the window opens (read before the window's commits),
    // plus, in a new-pixel segment, the window's own pixels filtered by rank
    pa.mask = j->mask.p;
    if (!redo) {
        const size_t tiles = (size_t)j->tiles_x * j->tiles_y;
        HIP_TRY(hipMemsetAsync(j->mend.p, 0, tiles * 8, s));
        HIP_TRY(hipMemsetAsync(j->tmin.p, 0xff, tiles * 4, s));
        HIP_TRY(launch_mask_or(j->rlist.p, (uint32_t)(j->locked + i0), (uint32_t)(j->locked + w1), j->W, j->tiles_x,
                               j->mend.p, j->tmin.p, s));
        pa.mask_new = j->mend.p;
        pa.tile_min = j->tmin.p;
        pa.rank = j->rank.p;
    }
    HIP_TRY(timed_launch(j, K_PREFETCH, n, [&] { return launch_knn_prefetch(pa, s); }));
    st.a.pre_nbrs = j->pre_nbrs.p;
    st.a.pre_cnt = j->pre_cnt.p;
    st.a.pre_base = (uint32_t)i0;
    st.a.pre_end = (uint32_t)w1;
    return TSG_OK;
}

// The batch [i0, i1) has been enqueued: advance the cursor and counts.
void step_advance(tsg_job* j, uint64_t i0, uint64_t i1) {
    tsg_job::Step& st = j->step;
    const uint64_t n_new = i1 > st.redo_eff ? i1 - std::max(i0, st.redo_eff) : 0;
    j->n_res += n_new;
    j->n_batches += 1;
    st.i = i1;
}

int step_close_stage(tsg_job* j) {
    tsg_job::Step& st = j->step;
    HIP_TRY(timed_launch(j, K_STATS, st.S_eff, [&] {
        return launch_stats_chunks(j->item_score.p, (uint32_t)st.S_eff, j->st_part.p, j->st_part_n.p,
                                   j->st_sum.p + st.t, j->st_cnt.p + st.t, j->ctx->stream);
    }));
    if (j->ahead_on) {
        tsg_job::Ahead& A = j->aset(st.t);
        if (!A.done) HIP_TRY(hipEventCreateWithFlags(&A.done, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(A.done, j->ctx->stream));
        A.done_recorded = true;
    }
    j->st_items[st.t] = st.S_eff;
    j->overall_cur += st.S;
    st.open = false;
    st.t += 1;
    return TSG_OK;
}

bool step_done(const tsg_job* j) { return j->step.t > (int)j->prm.backtrack_stages; }

// After the job's last launch: record the end event; step_finish waits.
// Side-stream work (pieces past the last item used) is joined first.
int step_end(tsg_job* j) {
    if (j->ahead_on) {
        hipEvent_t e;
        HIP_TRY(next_event(j, &e));
        HIP_TRY(hipEventRecord(e, j->ctx->side));
        HIP_TRY(hipStreamWaitEvent(j->ctx->stream, e, 0));
    }
    HIP_TRY(hipEventRecord(j->ev_run1, j->ctx->stream));
    return TSG_OK;
}

int step_finish(tsg_job* j) {
    HIP_TRY(hipEventSynchronize(j->ev_run1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, j->ev_run0, j->ev_run1));
    (void)hipEventDestroy(j->ev_run0);
    (void)hipEventDestroy(j->ev_run1);
    j->ev_run0 = j->ev_run1 = nullptr;
    j->run_ms = ms;
    std::vector<unsigned long long> spans(2 * j->launch_slots);
    if (j->launch_slots)
        HIP_TRY(hipMemcpy(spans.data(), j->launch_clk.p, spans.size() * 8, hipMemcpyDeviceToHost));
    for (const tsg_job::Timed& t : j->timed) {
        float km = 0.0f;
        if (t.a) HIP_TRY(hipEventElapsedTime(&km, t.a, t.b));
        if (t.slot >= 0 && spans[2 * t.slot] != 0) {
            const unsigned long long a0 = ~spans[2 * t.slot], a1 = spans[2 * t.slot + 1];
            if (a0 <= a1) km = (float)((double)(a1 - a0) * 1e-5);  // 100 MHz device clock, in ms
        }
        j->kern_ms[t.kind] += km;
        j->kern_items[t.kind] += t.items;
        j->kern_launches[t.kind] += 1;
    }
    uint32_t ovf = 0;
    HIP_TRY(hipMemcpy(&ovf, j->overflow.p, 4, hipMemcpyDeviceToHost));
    j->step.active = false;
    j->ran = true;
    if (ovf & 4u) return fail(TSG_ERR_DEVICE, "a dependency wait of a flow run timed out");
    if (ovf & 2u) return fail(TSG_ERR_DEVICE, "a batch exceeded its cooperative launch");
    if (ovf) return fail(TSG_ERR_SAMPLING, "a random-candidate rejection loop hit its bound (sample mask too sparse)");
    return TSG_OK;
}

// Single job: with stage-ahead neighbour lists, each stage's leading run of
// new-pixel batches goes to one dependency-ordered launch (k_resolve_flow);
// other runs of small batches go to one single-workgroup launch; other
// batches get one k_resolve launch, committing in its last workgroup when
// the batch holds at most fuse_max items.
int run_job(tsg_job* j, tsg_progress_fn cb, void* user) {
    hipStream_t s = j->ctx->stream;
    int rc;
    if ((rc = step_begin(j))) return rc;
    // content-independent work ahead on the side stream when the buffer
    // sets fit the budget: one per stage (deep), else two of the largest stage
    const bool deep_fits = j->prm.backtrack_stages + 1 <= 8 && ahead_bytes_deep(j) <= kAheadBudget;
    j->ahead_deep = j->tune.ahead >= 2 && !j->prm.tiling_mode && deep_fits;
    j->ahead_on = j->ahead_deep || (j->tune.ahead && !j->prm.tiling_mode && ahead_bytes(j) <= kAheadBudget);
    if (j->ahead_on) {
        const uint64_t smax = std::max<uint64_t>(1, stage_size(j, 0));
        for (int q = 0; q < 2 && !j->ahead_deep; ++q) {
            tsg_job::Ahead& A = j->ahead[q];
            HIP_TRY(A.rc.reserve(smax * j->prm.random_sample_locations));
            HIP_TRY(A.nbrs.reserve(smax * j->prm.nearest_neighbors));
            HIP_TRY(A.cnt.reserve(smax));
            HIP_TRY(A.mend.reserve((size_t)j->tiles_x * j->tiles_y));
            HIP_TRY(A.mnew.reserve((size_t)j->tiles_x * j->tiles_y));
            HIP_TRY(A.tmin.reserve((size_t)j->tiles_x * j->tiles_y));
            HIP_TRY(A.dep.reserve(smax));
            if (!A.ready) HIP_TRY(hipEventCreateWithFlags(&A.ready, hipEventDisableTiming));
        }
        j->ahead_open = j->n_res;
        j->ahead_next = 0;
        // the side stream starts after the set-up work on the job stream
        hipEvent_t e;
        HIP_TRY(next_event(j, &e));
        HIP_TRY(hipEventRecord(e, s));
        HIP_TRY(hipStreamWaitEvent(j->ctx->side, e, 0));
        if ((rc = ahead_stage(j))) return rc;
    }
    // dependency-ordered runs need the stage-ahead neighbour lists
    bool flow_on = j->ahead_on && j->tune.flow;
    const uint32_t flow_team = j->tune.flow_team;
    uint32_t flow_wgs = 0;
    if (flow_on) {
        const uint32_t mx = flow_max_wgs(flow_team, j->guided);
        flow_wgs = j->tune.flow_wgs ? std::min(j->tune.flow_wgs, mx) : mx;
        flow_on = flow_wgs > 0;
    }
    if (flow_on) {
        if (j->flow_flag.n < j->N) {
            HIP_TRY(j->flow_flag.alloc(j->N));
            HIP_TRY(hipMemsetAsync(j->flow_flag.p, 0, j->N * 4, s));
            j->flow_epoch = 0;
        }
        if (j->tune.flow >= 2 && j->state_old.n < j->N) HIP_TRY(j->state_old.alloc(j->N));
    }
    // persistent single-wave batches
    const uint32_t persist_wgs = j->tune.persist ? persist_max_wgs(j->guided) : 0u;
    if (flow_on && !j->flow_ticket.n) {
        // the ticket counter of dependency-ordered runs
        HIP_TRY(j->flow_ticket.alloc(1));
        HIP_TRY(hipMemsetAsync(j->flow_ticket.p, 0, 4, s));
        j->flow_tickets = 0;
    }
    // deferred commits for throughput-bound batches (not in tiling mode,
    // whose commits also set ghost flags)
    const bool defer_on = j->tune.defer && !j->prm.tiling_mode;
    const size_t tiles = (size_t)j->tiles_x * j->tiles_y;
    if (defer_on) {
        if (j->state2.n < j->N) {
            HIP_TRY(j->state2.alloc(j->N));
            HIP_TRY(j->stamp.alloc(j->N));
            HIP_TRY(hipMemsetAsync(j->stamp.p, 0, j->N * 4, s));
            j->serial = 2;
        }
        if (j->pend.n < 3 * tiles) {
            HIP_TRY(j->pend.alloc(3 * tiles));
            HIP_TRY(hipMemsetAsync(j->pend.p, 0, 3 * tiles * 8, s));
        }
    }
    // the batch whose commit is pending [dp0, dp1), and the one before it,
    // whose pending bits are still set [dq0, dq1)
    bool d_pending = false;
    uint64_t dp0 = 0, dp1 = 0, dq0 = 0, dq1 = 0;
    auto pend_at = [&](uint32_t sr) { return j->pend.p + (size_t)(sr % 3) * tiles; };
    auto flush_def = [&]() -> int {
        if (!d_pending) return TSG_OK;
        const DeferFlushArgs fa{j->W,       j->tiles_x,          j->step.a.work,          j->state2.p,
                                j->state.p, j->mask.p,           pend_at(j->serial),      pend_at(j->serial - 1),
                                j->step.a.first_new, (uint32_t)dp0, (uint32_t)dp1, (uint32_t)dq0, (uint32_t)dq1};
        HIP_TRY(timed_launch(j, K_COMMIT, dp1 - dp0, [&] { return launch_defer_flush(fa, s); }));
        d_pending = false;
        dq0 = dq1 = 0;
        j->serial += 2;  // no stamp names the batch before the next one
        return TSG_OK;
    };
    int last_pct = 0;
    std::vector<uint8_t> snap;
    auto progress = [&](uint64_t i1) -> int {
        if (int rc3 = flush_def()) return rc3;
        int pct = (int)std::lround((double)(j->overall_cur + i1) / (double)std::max<uint64_t>(j->overall_total, 1) *
                                   100.0);
        if (pct == last_pct) return TSG_OK;
        last_pct = pct;
        snap.resize(j->N * 4);
        HIP_TRY(hipStreamSynchronize(s));
        std::vector<uint4> stv(j->N);
        HIP_TRY(hipMemcpy(stv.data(), j->state.p, j->N * sizeof(uint4), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < j->N; ++q) memcpy(&snap[q * 4], &stv[q].x, 4);
        cb(user, snap.data(), j->W, j->H, j->overall_cur + i1, j->overall_total, i1, j->step.S);
        return TSG_OK;
    };
    // items [i0, end) of the open stage in one dependency-ordered launch;
    // old: a redo run, which reads neighbours not redone before it from the
    // state as the segment opened
    auto flow_run = [&](uint64_t i0, uint64_t end, const uint4* old) -> int {
        if (int rc3 = flush_def()) return rc3;
        tsg_job::Step& st = j->step;
        BatchArgs& a = st.a;
        a.item0 = (uint32_t)i0;
        a.item1 = (uint32_t)end;
        a.n_res = (uint32_t)j->n_res;
        if (++j->flow_epoch == 0) {
            HIP_TRY(hipMemsetAsync(j->flow_flag.p, 0, j->N * 4, s));
            j->flow_epoch = 1;
        }
        // the run's leading items (its longest serial chain) in a launch of
        // their own with larger teams; the rest follow in stream order, under
        // the same epoch, so their waits on the head are already satisfied
        uint64_t h = 0;
        const uint32_t head_wgs = j->tune.flow_head ? flow_max_wgs(j->tune.flow_head_team, j->guided) : 0u;
        if (head_wgs && i0 + j->tune.flow_head < end) {
            h = j->tune.flow_head;
            const uint32_t nh = (uint32_t)h;
            const uint32_t gh = std::min(head_wgs, nh);
            const FlowArgs fh{j->flow_ticket.p, j->flow_tickets, nh, j->aset(st.t).dep.p, j->flow_flag.p,
                              j->flow_epoch, kFlowSpin, j->overflow.p, old};
            HIP_TRY(timed_launch(j, K_RUN, nh, [&] {
                BatchArgs fa = a;
                fa.item1 = (uint32_t)(i0 + h);
                fa.result = nullptr;
                return launch_resolve_flow(fa, fh, gh, j->tune.flow_head_team, j->guided, s);
            }));
            j->flow_tickets += nh + gh;
        }
        const uint32_t n = (uint32_t)(end - i0 - h);
        const uint32_t g = std::min(flow_wgs, n);
        const FlowArgs f{j->flow_ticket.p, j->flow_tickets, n, j->aset(st.t).dep.p, j->flow_flag.p,
                         j->flow_epoch, kFlowSpin, j->overflow.p, old};
        HIP_TRY(timed_launch(j, K_RUN, n, [&] {
            BatchArgs fa = a;  // results stay in registers
            fa.item0 = (uint32_t)(i0 + h);
            fa.result = nullptr;
            return launch_resolve_flow(fa, f, g, flow_team, j->guided, s);
        }));
        j->flow_tickets += n + g;
        while (st.i < end) {
            const uint64_t b0 = st.i, b1 = batch_end(j, b0, st.redo, st.S_eff);
            step_advance(j, b0, b1);
        }
        if (st.i != end) return fail(TSG_ERR_DEVICE, "dependency-ordered run out of step");
        if (cb) return progress(end);
        return TSG_OK;
    };
    std::vector<uint32_t> sb_bounds, sb_nres;
    bool sb_coop = false;  // the pending run goes to the cooperative kernel
    const uint32_t coop_wgs = std::min(j->tune.coop_wgs, j->tune.coop_wgs ? coop_max_wgs(j->guided) : 0u);
    const uint64_t coop_max = std::min<uint64_t>(j->tune.coop_max, (uint64_t)coop_wgs * small_batch_max());
    auto flush_small = [&]() -> int {
        if (sb_nres.empty()) return TSG_OK;
        int rc2 = flush_def();
        if (rc2) return rc2;
        rc2 = step_ensure_rc(j, sb_bounds.front(), sb_bounds.back());
        if (rc2) return rc2;
        std::vector<uint32_t> tab(sb_bounds);
        tab.insert(tab.end(), sb_nres.begin(), sb_nres.end());
        const size_t nb = sb_nres.size();
        // the table lives in a per-run device arena: bounds (nb+1) then nres (nb)
        if (j->small_used + tab.size() > j->small_tab.n) return fail(TSG_ERR_DEVICE, "small-batch table overflow");
        uint32_t* dt = j->small_tab.p + j->small_used;
        j->small_used += tab.size();
        HIP_TRY(j->ctx->staging.upload(dt, tab.data(), tab.size() * 4, s));
        const BatchArgs& a = j->step.a;
        if (sb_coop) {
            HIP_TRY(timed_launch(j, K_RUN, sb_bounds.back() - sb_bounds.front(), [&] {
                return launch_resolve_coop(a, dt, dt + nb + 1, (uint32_t)nb, coop_wgs,
                                           j->tune.team_max >= 8 ? 16u : j->tune.team_max, j->overflow.p,
                                           j->guided, s);
            }));
        } else {
            HIP_TRY(timed_launch(j, K_RUN, sb_bounds.back() - sb_bounds.front(), [&] {
                return launch_resolve_small(a, dt, dt + nb + 1, (uint32_t)nb, j->guided, s);
            }));
        }
        sb_bounds.clear();
        sb_nres.clear();
        return TSG_OK;
    };
    while (!step_done(j)) {
        if ((rc = step_open_stage(j))) return rc;
        uint64_t i0, i1;
        while (step_peek(j, i0, i1)) {
            if ((rc = ahead_poll(j, false))) return rc;
            tsg_job::Step& st = j->step;
            const uint64_t flow_end = flow_on ? std::min<uint64_t>(j->aset(st.t).w_end, st.S_eff) : 0;
            if (flow_on && j->tune.flow >= 2 && i0 == 0 && st.redo_eff > 0 && flow_end > 0) {
                // the stage's leading redo batches: one run over a snapshot
                // of the resolved pixels' state
                if ((rc = flush_small())) return rc;
                HIP_TRY(launch_state_snapshot(j->rlist.p, (uint32_t)j->n_res, j->state.p, j->state_old.p, s));
                if ((rc = flow_run(0, std::min(flow_end, st.redo_eff), j->state_old.p))) return rc;
                continue;
            }
            if (flow_on && i0 == st.redo_eff && flow_end > i0) {
                // the stage's new-pixel batches with neighbour lists: one
                // dependency-ordered run
                if ((rc = flush_small())) return rc;
                if ((rc = flow_run(i0, flow_end, nullptr))) return rc;
                continue;
            }
            if (step_wants_prefetch(j, i0, i1)) {
                // earlier small batches read the previous window: launch them first
                if ((rc = flush_small())) return rc;
                if ((rc = flush_def())) return rc;
                if ((rc = step_prefetch(j, i0))) return rc;
            }
            const bool coop = coop_wgs > 0 && i1 - i0 <= coop_max;
            if (coop || i1 - i0 <= j->tune.small_max) {
                if (!sb_nres.empty() && sb_coop != coop && (rc = flush_small())) return rc;
                sb_coop = coop;
                if (sb_bounds.empty()) sb_bounds.push_back((uint32_t)i0);
                sb_bounds.push_back((uint32_t)i1);
                sb_nres.push_back((uint32_t)j->n_res);
            } else {
                if ((rc = flush_small())) return rc;
                if ((rc = step_ensure_rc(j, i0, i1))) return rc;
                BatchArgs& a = j->step.a;
                a.n_res = (uint32_t)j->n_res;
                a.item0 = (uint32_t)i0;
                a.item1 = (uint32_t)i1;
                a.fused_commit = (i1 - i0) <= j->tune.fuse_max;
                a.team = team_for(i1 - i0, j->tune.team_max, j->tune.team_waves);
                if (defer_on && !a.fused_commit) {
                    // commits the pending batch; its own commit is left to the next launch
                    const uint32_t sr = ++j->serial;
                    BatchArgs da = a;
                    da.result = nullptr;
                    da.state2 = j->state2.p;
                    da.stamp = j->stamp.p;
                    da.serial = sr;
                    da.pend_r = pend_at(sr - 1);
                    da.pend_w = pend_at(sr);
                    da.pend_c = pend_at(sr - 2);
                    if (d_pending) {
                        da.prev0 = (uint32_t)dp0;
                        da.prev1 = (uint32_t)dp1;
                        da.pp0 = (uint32_t)dq0;
                        da.pp1 = (uint32_t)dq1;
                    }
                    HIP_TRY(timed_launch(j, K_RESOLVE, i1 - i0, [&] {
                        da.dbg_launch = a.dbg_launch;  // set by the profiler for this launch
                        return launch_resolve(da, j->guided, s);
                    }));
                    dq0 = d_pending ? dp0 : 0;
                    dq1 = d_pending ? dp1 : 0;
                    dp0 = i0;
                    dp1 = i1;
                    d_pending = true;
                    step_advance(j, i0, i1);
                    if (cb && (rc = progress(i1))) return rc;
                    continue;
                }
                if ((rc = flush_def())) return rc;
                if (persist_wgs && a.team <= 1 && !a.fused_commit) {
                    const uint32_t g = (uint32_t)std::min<uint64_t>(persist_wgs, (i1 - i0 + 7) / 8);
                    HIP_TRY(timed_launch(j, K_RESOLVE, i1 - i0,
                                         [&] { return launch_resolve_persist(a, g, j->guided, s); }));
                } else {
                    HIP_TRY(timed_launch(j, K_RESOLVE, i1 - i0, [&] { return launch_resolve(a, j->guided, s); }));
                }
                if (!a.fused_commit) {
                    CommitArgs ca{j->W,       j->tiles_x, a.maps,    a.work,      j->result.p, (uint32_t)i0, (uint32_t)i1,
                                  a.first_new, j->state.p, j->mask.p, j->score.p, a.gmask};
                    HIP_TRY(timed_launch(j, K_COMMIT, i1 - i0, [&] { return launch_commit(ca, s); }));
                }
            }
            step_advance(j, i0, i1);
            if (cb) {
                if ((rc = flush_small())) return rc;
                if ((rc = progress(i1))) return rc;
            }
        }
        if ((rc = flush_small())) return rc;
        if ((rc = flush_def())) return rc;
        // the next stage's set must be prepared before that stage opens
        if ((rc = ahead_poll(j, true))) return rc;
        if ((rc = step_close_stage(j))) return rc;
    }
    if ((rc = step_end(j))) return rc;
    return step_finish(j);
}

// Launch structures the autotuner stores apply to jobs of the same shape:
// output size, k, m, sampled maps, guided, tiling, stages.
std::vector<uint64_t> tuning_key(const tsg_job* j) {
    return {j->W, j->H, j->prm.nearest_neighbors, j->prm.random_sample_locations, j->filt.size(),
            (uint64_t)j->guided, j->prm.tiling_mode, j->prm.backtrack_stages};
}

int check_tuning(const tsg_tuning* t) {
    if (t->small_max > small_batch_max()) return fail(TSG_ERR_INVALID_ARGUMENT, "small_max above 16");
    if (t->team_max == 0 || t->team_max > 8 || (t->team_max & (t->team_max - 1)))
        return fail(TSG_ERR_INVALID_ARGUMENT, "team_max must be 1, 2, 4 or 8");
    if (t->coop_max > t->coop_wgs * small_batch_max())
        return fail(TSG_ERR_INVALID_ARGUMENT, "coop_max above 16 items per cooperative workgroup");
    if (t->flow > 2) return fail(TSG_ERR_INVALID_ARGUMENT, "flow must be 0, 1 or 2");
    if (t->defer > 1) return fail(TSG_ERR_INVALID_ARGUMENT, "defer must be 0 or 1");
    if (t->persist > 1) return fail(TSG_ERR_INVALID_ARGUMENT, "persist must be 0 or 1");
    if (t->flow_head && (t->flow_head_team == 0 || t->flow_head_team > 8 ||
                         (t->flow_head_team & (t->flow_head_team - 1))))
        return fail(TSG_ERR_INVALID_ARGUMENT, "flow_head_team must be 1, 2, 4 or 8");
    if (t->flow && (t->flow_team == 0 || t->flow_team > 8 || (t->flow_team & (t->flow_team - 1))))
        return fail(TSG_ERR_INVALID_ARGUMENT, "flow_team must be 1, 2, 4 or 8");
    if (t->ahead > 2) return fail(TSG_ERR_INVALID_ARGUMENT, "ahead must be 0, 1 or 2");
    return TSG_OK;
}

// ---- multi-job engine ----

size_t job_footprint(const tsg_job* j) {
    size_t b = 0;
    for (auto& e : j->ex) b += e->n * 4;
    for (auto& e : j->exg) b += e->n * 4;
    for (auto& e : j->smask) b += e->n;
    b += j->tg.n * 4 + j->maps.n * sizeof(MapDesc);
    b += j->N * (sizeof(uint4) + 4 + 4 + 4) + (size_t)j->tiles_x * j->tiles_y * 8 + j->gmask.n * 8;
    b += j->result.n * sizeof(uint4) + j->rc.n * sizeof(uint2) + j->small_tab.n * 4;
    b += j->rank.n * 4 + j->mend.n * 8 + j->tmin.n * 4 + j->pre_n.n * 4 + j->pre_cnt.n * 4 + j->pre_nbrs.n * 4;
    for (const tsg_job::Ahead& A : j->ahead)
        b += A.rc.n * sizeof(uint2) + A.nbrs.n * 4 + A.cnt.n * 4 + A.n_at.n * 4 + A.mend.n * 8 + A.mnew.n * 8 +
             A.tmin.n * 4 + A.dep.n * 8;
    b += j->flow_flag.n * 4 + j->flow_ticket.n * 4 + j->state_old.n * sizeof(uint4);
    b += j->state2.n * sizeof(uint4) + j->stamp.n * 4 + j->pend.n * 8;
    b += j->st_part.n * 8 + j->st_part_n.n * 8 + j->eng.luts.n * 4;
    return b;
}

}  // namespace

struct tsg_batch {
    tsg_ctx* ctx = nullptr;
    uint32_t max_active = 0;
    struct Entry {
        uint64_t ticket;
        tsg_job* job;
        int status;
        bool failed;  // status set by the engine (planning error, cancellation)
        bool alone;   // runs (ran) alone on the single-job path; status final once run
    };
    std::vector<Entry> active, pending, done, taken;
    uint64_t next_ticket = 1;
    // one launch's tables, built in pinned memory and copied in one piece:
    // [QCtl x 2 | QJob | QDyn | QPhase | BatchArgs | n_at]
    char* h_blk = nullptr;
    size_t h_cap = 0;
    DBuf<char> d_blk;
    uint32_t wgs[2] = {0, 0};  // resident engine workgroups (unguided, guided)
    uint64_t rounds = 0, launches = 0, units_total = 0;
    uint64_t idle_ticks = 200000000ull;  // 2 s of the 100 MHz clock without progress on an XCD
    uint32_t wgs_cap = 0;                 // test hook: at most this many engine workgroups (0: all resident)
    double prof[16] = {};                 // accumulated QXcd clocks and counts, host phase times (tsg_batch_engine_profile)
    uint32_t engine_flags = 0;            // test hook: kEngineSpread forces spread launches
    uint64_t quantum = 0;                 // resolve items per job per launch (0: no bound)
    uint32_t alone_below = 0;             // jobs starting while fewer are in the batch run alone (0: never)
    uint64_t alone_ahead = 0;             // device bytes an alone run may add for its stage-ahead sets
    uint64_t alone_runs = 0;              // jobs run alone so far
    tsg_tuning tune{};
    ~tsg_batch() {
        if (h_blk) (void)hipHostFree(h_blk);
    }
};

namespace {

// Phases of at most this many items commit in the batch's closing workgroup.
constexpr uint32_t kEngineFuseMax = 2048;
// Launches with at most this many jobs compute the neighbour lists of
// latency-bound batches ahead, as a single-job run does.
constexpr uint32_t kEnginePrefetchJobs = 8;
// Launches of fewer jobs than this are spread over every XCD (home placement
// would give a lone job one XCD in eight).
constexpr size_t kSpreadBelow = 8;
// Items per team in one resolve unit when more jobs than that share a launch
// (256-job mix: 1, 2, 4, 8, 16 measured 1.22, 1.10, 1.02, 1.02, 1.10 s; the
// waves of a unit wait for its slowest item, and larger units average that
// out until the last units of the launch leave workgroups idle).
constexpr uint32_t kEngineRep = 4;
constexpr uint32_t kNoArg = 0xffffffffu;

void eng_push(tsg_job* j, uint32_t kind, uint32_t i0, uint32_t i1, uint32_t per_unit, uint32_t arg) {
    QPhase q{};
    q.kind = kind;
    q.stage = (uint32_t)j->step.t;
    q.i0 = i0;
    q.i1 = i1;
    q.units = (i1 - i0 + per_unit - 1) / per_unit;
    q.arg = arg == kNoArg ? j->eng.stage_arg : arg;
    if (q.units) j->eng.prog.push_back(q);
}

// Stage preamble (ms.rs:786-858) as planned work: the recolour phase, the
// stage's cost tables and picks uploaded, the stage's batch arguments.
int eng_open_stage(tsg_job* j) {
    hipStream_t s = j->ctx->stream;
    tsg_job::Step& st = j->step;
    const int P = (int)j->prm.backtrack_stages;
    const int t = st.t;
    const int ps = P - t;
    const uint32_t nf = (uint32_t)j->filt.size();
    st.level = t == 0 ? 0u : (uint32_t)std::min(t, P - 1);
    const MapDesc* maps = j->maps.p + (size_t)st.level * nf;
    st.pss = (uint64_t)Pcg32::seed_from_u64(j->prm.seed + (uint64_t)ps).next_u32();
    st.S = stage_size(j, ps);
    st.redo = j->n_res - j->locked;
    float alpha_s = 0.0f;
    if (j->guided && ps > 0) {
        float b = j->prm.guide_alpha * (1.0f - ((float)j->n_res / (float)j->N0));
        alpha_s = (b * b) * b;
    }
    stage_luts(j->prm.cauchy_dispersion, j->guided, alpha_s, j->lut_keep[t].data());
    uint32_t lut_nan = 0;
    for (float v : j->lut_keep[t]) lut_nan |= (uint32_t)std::isnan(v);
    float* lut = j->eng.luts.p + (size_t)t * 512;
    HIP_TRY(j->ctx->staging.upload(lut, j->lut_keep[t].data(), 512 * 4, s));
    int rc = wait_picks(j, t);
    if (rc) return rc;
    st.S_eff = st.S;
    if (st.S > st.redo) {
        st.S_eff = st.redo + j->picks_keep[t].size();
        if (!j->picks_keep[t].empty()) {
            HIP_TRY(j->ctx->staging.upload(j->rlist.p + j->locked + st.redo, j->picks_keep[t].data(),
                                           j->picks_keep[t].size() * 4, s));
            const uint32_t lo = (uint32_t)(j->locked + st.redo);
            HIP_TRY(launch_rank(j->rlist.p, lo, lo + (uint32_t)j->picks_keep[t].size(), j->rank.p, s));
        }
    }
    st.redo_eff = std::min(st.redo, st.S_eff);
    st.i = 0;
    st.rc_lo = st.rc_hi = 0;
    BatchArgs& a = st.a;
    a = BatchArgs{};
    a.W = j->W;
    a.H = j->H;
    a.tiles_x = j->tiles_x;
    a.tiles_y = j->tiles_y;
    a.mask = j->mask.p;
    a.state = j->state.p;
    a.tg = j->guided ? j->tg.p + (size_t)st.level * j->N : nullptr;
    a.maps = maps;
    a.n_maps = nf;
    a.work = j->rlist.p + j->locked;
    a.rlist = j->rlist.p;
    a.m = (uint32_t)j->prm.random_sample_locations;
    a.offtab = j->ctx->offtab.p;
    a.offtab_n = j->ctx->offtab_n;
    a.lut = lut;
    a.stage_seed = st.pss;
    a.k = j->prm.nearest_neighbors;
    a.item_score = j->item_score.p;
    a.lut_nan = lut_nan;
    a.first_new = (uint32_t)st.redo_eff;
    a.state_w = j->state.p;
    a.mask_w = j->mask.p;
    a.score_w = j->score.p;
    a.result = j->result.p;
    a.rcands = j->rc.p;
    if (j->prm.tiling_mode) {
        a.gmask = j->gmask.p;
        a.xl = (int32_t)((float)j->W * 0.05f);
        a.xr = (int32_t)j->W - a.xl;
        a.yb = (int32_t)((float)j->H * 0.05f);
        a.yt = (int32_t)j->H - a.yb;
    }
    j->eng.stage_arg = (uint32_t)j->eng.args.size();
    j->eng.args.push_back(a);
    j->eng.stage_a = a;
    // next_pyramid_level (ms.rs:687-700): every resolved pixel from the new level
    if (st.level > 0) eng_push(j, Q_RECOLOR, 0, (uint32_t)j->n_res, kQElems, kNoArg);
    st.open = true;
    return TSG_OK;
}

// The prefetch window that starts at i0 (as step_prefetch plans it), as
// phases: the window's new pixels into the window masks, then one wave per
// item's neighbour list.
void eng_prefetch(tsg_job* j, uint64_t i0) {
    tsg_job::Step& st = j->step;
    const bool redo = i0 < st.redo_eff;
    const uint64_t seg_end = redo ? st.redo_eff : st.S_eff;
    const uint32_t off = (uint32_t)j->eng.n_at.size();
    uint64_t i = i0, nres = j->n_res;
    while (i < seg_end) {
        const uint64_t e = batch_end_at(j, st.t, i, nres, st.redo, st.S_eff);
        if (e - i > j->tune.prefetch_max || e - i0 > kPrefetchCap) break;
        j->eng.n_at.insert(j->eng.n_at.end(), e - i, (uint32_t)nres);
        if (e > st.redo_eff) nres += e - std::max(i, st.redo_eff);
        i = e;
    }
    const uint64_t w1 = i;
    if (w1 <= i0) return;
    if (!redo) {
        eng_push(j, Q_CLEAR, 0, j->tiles_x * j->tiles_y, kQElems, kNoArg);
        eng_push(j, Q_MASKOR, (uint32_t)(j->locked + i0), (uint32_t)(j->locked + w1), kQElems, kNoArg);
    }
    eng_push(j, Q_PREFETCH, (uint32_t)i0, (uint32_t)w1, 8, kNoArg);
    j->eng.prog.back().pre_base = off;
    j->eng.prog.back().pre_end = redo ? 0u : 1u;
    st.a.pre_nbrs = j->pre_nbrs.p;
    st.a.pre_cnt = j->pre_cnt.p;
    st.a.pre_base = (uint32_t)i0;
    st.a.pre_end = (uint32_t)w1;
}

// Plans the job as far as its pick orders allow: the same batches, random
// candidate chunks and prefetch windows as a run alone, as phases of the
// next launch.
int eng_plan(tsg_job* j) {
    tsg_job::Step& st = j->step;
    int rc;
    // a stage left open by the last launch (its quantum): this launch's
    // phases of the stage need the stage's block among this launch's blocks
    if (st.open && j->eng.args.empty()) {
        j->eng.stage_arg = 0;
        j->eng.args.push_back(j->eng.stage_a);
    }
    while (!step_done(j)) {
        if (!st.open) {
            if (!picks_ready_now(j, st.t)) return TSG_OK;
            if ((rc = eng_open_stage(j))) return rc;
        }
        uint64_t i0 = 0, i1 = 0;
        if (!step_peek(j, i0, i1)) {
            // stage statistics, then the stage is through (step_close_stage)
            eng_push(j, Q_STATS, 0, (uint32_t)st.S_eff, kStatsChunk, kNoArg);
            j->st_items[st.t] = st.S_eff;
            j->overall_cur += st.S;
            st.open = false;
            st.t += 1;
            continue;
        }
        // the launch's share of this job is planned (at least one batch)
        if (j->eng.quantum && j->eng.items >= j->eng.quantum) return TSG_OK;
        const uint32_t m = (uint32_t)j->prm.random_sample_locations;
        if (!(i0 >= st.rc_lo && i1 <= st.rc_hi)) {
            const uint64_t cap_items = std::max<uint64_t>(i1 - i0, (uint64_t)j->rc_budget / (8ull * std::max(1u, m)));
            st.rc_lo = i0;
            st.rc_hi = std::min<uint64_t>(st.S_eff, i0 + cap_items);
            if ((st.rc_hi - st.rc_lo) * m > j->rc.n)
                return fail(TSG_ERR_DEVICE, "random-candidate chunk beyond its buffer");
            eng_push(j, Q_RAND, (uint32_t)st.rc_lo, (uint32_t)st.rc_hi, kQElems, kNoArg);
        }
        // neighbour lists ahead of latency-bound batches shorten a lone job's
        // chain; with other jobs to run meanwhile the resolve searches itself
        if (j->eng.sharing <= kEnginePrefetchJobs && step_wants_prefetch(j, i0, i1)) eng_prefetch(j, i0);
        // teams of waves per item shorten a batch's latency at the cost of
        // throughput; the jobs of a launch share the device, so the team is
        // sized against the waves per job
        const uint32_t team = team_for((i1 - i0) * j->eng.sharing, j->tune.team_max, j->tune.team_waves);
        const uint32_t n = (uint32_t)(i1 - i0);
        // the batch's own arguments ride on its phase; the rest are the stage's
        // with many jobs sharing the device, a unit takes several items per
        // team: fewer claims per item, and parallelism enough in other jobs
        const uint32_t rep = j->eng.sharing > kEnginePrefetchJobs ? kEngineRep : 1u;
        eng_push(j, Q_RESOLVE, (uint32_t)i0, (uint32_t)i1, items_per_wg(team) * rep, kNoArg);
        QPhase& q = j->eng.prog.back();
        q.rep = rep;
        q.n_res = (uint32_t)j->n_res;
        q.team = team;
        q.rc_base = (uint32_t)st.rc_lo;
        if (st.a.pre_nbrs && i0 >= st.a.pre_base && i1 <= st.a.pre_end) {
            q.pre_base = st.a.pre_base;
            q.pre_end = st.a.pre_end;
        }
        q.fuse = n <= kEngineFuseMax;
        if (!q.fuse) eng_push(j, Q_COMMIT, (uint32_t)i0, (uint32_t)i1, kQElems, kNoArg);
        j->eng.weight += (uint64_t)n * (j->prm.nearest_neighbors + 8);
        j->eng.items += n;
        step_advance(j, i0, i1);
    }
    j->eng.planned_all = true;
    return TSG_OK;
}

// Engine buffers of an admitted job, sized for its largest batch, chunk and window.
int eng_begin(tsg_job* j) {
    int rc = step_begin(j);
    if (rc) return rc;
    const int P = (int)j->prm.backtrack_stages;
    HIP_TRY(j->eng.luts.reserve((size_t)(P + 1) * 512));
    const uint64_t m = std::max<uint64_t>(1, j->prm.random_sample_locations);
    const uint64_t maxb = j->result.n;
    HIP_TRY(j->rc.reserve(std::min<uint64_t>(j->N, std::max<uint64_t>(maxb, j->rc_budget / (8 * m))) * m));
    const uint64_t win = std::min<uint64_t>(kPrefetchCap, j->N);
    HIP_TRY(j->pre_cnt.reserve(win));
    HIP_TRY(j->pre_nbrs.reserve(win * j->prm.nearest_neighbors));
    const size_t tiles = (size_t)j->tiles_x * j->tiles_y;
    HIP_TRY(j->mend.reserve(tiles));
    HIP_TRY(j->tmin.reserve(tiles));
    j->eng.prog.clear();
    j->eng.args.clear();
    j->eng.n_at.clear();
    j->eng.stage_arg = 0;
    j->eng.weight = 0;
    j->eng.items = 0;
    j->eng.planned_all = false;
    return TSG_OK;
}

// A job leaves the engine with a status of its own; its queued uploads (if
// any) finish before the end event.
int eng_drop(tsg_batch* b, size_t q, int status) {
    tsg_batch::Entry e = b->active[q];
    b->active.erase(b->active.begin() + q);
    e.status = status;
    e.failed = true;
    tsg_job* j = e.job;
    if (j->ev_run1) (void)hipEventRecord(j->ev_run1, b->ctx->stream);
    j->eng.prog.clear();
    j->eng.args.clear();
    j->eng.n_at.clear();
    b->done.push_back(e);
    return TSG_OK;
}

int batch_admit(tsg_batch* b) {
    while (b->active.size() < b->max_active && !b->pending.empty()) {
        tsg_batch::Entry e = b->pending.front();
        b->pending.erase(b->pending.begin());
        // few jobs in the batch: this one gets the whole device to itself
        e.alone = b->alone_below && b->active.size() + b->pending.size() + 1 < b->alone_below;
        b->active.push_back(e);
        if (e.alone) continue;
        const int rc = eng_begin(e.job);
        if (rc) eng_drop(b, b->active.size() - 1, rc);
    }
    return TSG_OK;
}

// Jobs admitted to run alone: each runs to its end on the single-job path
// (tsg_job_run's launch sequence), one after another, smallest first, before
// the round's engine launch. Its stage-ahead sets (which the dependency-
// ordered runs need) are used only when they fit the bytes the caller allows
// beyond tsg_footprint_estimate, and are released when the run ends.
int batch_run_alone(tsg_batch* b) {
    std::vector<tsg_batch::Entry> runs;
    for (size_t q = 0; q < b->active.size();) {
        if (!b->active[q].alone) {
            ++q;
            continue;
        }
        runs.push_back(b->active[q]);
        b->active.erase(b->active.begin() + q);
    }
    std::stable_sort(runs.begin(), runs.end(),
                     [](const tsg_batch::Entry& x, const tsg_batch::Entry& y) { return x.job->N < y.job->N; });
    for (tsg_batch::Entry& e : runs) {
        tsg_job* j = e.job;
        // a single-job run (its small-batch tables, all pick threads); the
        // random candidates keep the engine's chunk size of the estimate
        j->in_engine = false;
        j->pk_threads = kPickThreads;
        if (j->tune.ahead >= 2 && ahead_bytes_deep(j) > b->alone_ahead) j->tune.ahead = 1;
        if (ahead_bytes(j) > b->alone_ahead) j->tune.ahead = 0;
        e.status = run_job(j, nullptr, nullptr);
        b->alone_runs += 1;
        // the run has ended (run_job waited for it, and the side stream is
        // drained here): its stage-ahead sets go back to the device cache
        // now, not with the job
        HIP_TRY(hipStreamSynchronize(j->ctx->side));
        g_pool_release = true;
        for (tsg_job::Ahead& A : j->ahead) {
            A.rc.release();
            A.nbrs.release();
            A.cnt.release();
            A.n_at.release();
            A.mend.release();
            A.mnew.release();
            A.tmin.release();
            A.dep.release();
        }
        g_pool_release = false;
        b->done.push_back(e);
    }
    return TSG_OK;
}

size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

// One launch of the engine over everything planned: jobs are placed on the
// XCDs by work (largest first onto the least loaded), their programs and
// batch arguments go to the device in one copy, then the unguided and the
// guided engine kernels run.
double host_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int batch_round(tsg_batch* b) {
    hipStream_t s = b->ctx->stream;
    double t_mark = host_ms();
    auto lap = [&](int slot) {
        const double t = host_ms();
        b->prof[slot] += t - t_mark;
        t_mark = t;
    };
    batch_admit(b);
    batch_run_alone(b);
    lap(6);
    if (b->active.empty()) return TSG_OK;
    // plan every job as far as its picks allow; a job that cannot be planned leaves
    for (size_t q = 0; q < b->active.size();) {
        b->active[q].job->eng.sharing = (uint32_t)b->active.size();
        b->active[q].job->eng.quantum = b->quantum;
        const int rc = eng_plan(b->active[q].job);
        if (rc) {
            eng_drop(b, q, rc);
            continue;
        }
        ++q;
    }
    // every phase's arguments must be among its job's blocks of this launch
    // (a planning error leaves with its job, before anything is launched)
    for (size_t q = 0; q < b->active.size();) {
        const tsg_job::Eng& E = b->active[q].job->eng;
        bool ok = true;
        for (const QPhase& p0 : E.prog) ok = ok && p0.arg < E.args.size();
        if (!ok) {
            eng_drop(b, q, fail(TSG_ERR_DEVICE, "engine phase without its arguments"));
            continue;
        }
        ++q;
    }
    std::vector<size_t> runs;
    for (size_t q = 0; q < b->active.size(); ++q)
        if (!b->active[q].job->eng.prog.empty()) runs.push_back(q);
    lap(7);
    if (runs.empty()) {
        // nothing planned: wait for one job's next pick order
        for (const tsg_batch::Entry& e : b->active)
            if (!e.job->eng.planned_all) {
                const int rc = wait_picks(e.job, e.job->step.t);
                if (rc) return rc;
                break;
            }
    } else {
        // placement: per kind (unguided 0, guided 1), largest work first onto the least loaded XCD
        std::vector<size_t> order[2];
        for (size_t q : runs) order[b->active[q].job->guided ? 1 : 0].push_back(q);
        std::vector<uint32_t> xcc_of(b->active.size(), 0);
        bool spread[2] = {false, false};
        for (int kd = 0; kd < 2; ++kd) {
            std::stable_sort(order[kd].begin(), order[kd].end(), [&](size_t x, size_t y) {
                return b->active[x].job->eng.weight > b->active[y].job->eng.weight;
            });
            // XCDs with engine workgroups: workgroup w runs on XCD w % 8
            if (!b->wgs[kd]) b->wgs[kd] = engine_max_wgs(kd == 1);
            const uint32_t wgs_kd = b->wgs_cap ? std::min(b->wgs_cap, b->wgs[kd]) : b->wgs[kd];
            const int n_x = (int)std::min<uint32_t>(8, std::max<uint32_t>(1, wgs_kd));
            // fewer jobs than XCDs: every XCD serves every job (a spread launch)
            spread[kd] = (b->engine_flags & kEngineSpread) || order[kd].size() < (size_t)kSpreadBelow;
            uint64_t load[8] = {};
            for (size_t q : order[kd]) {
                const int x = spread[kd] ? 0 : (int)(std::min_element(load, load + n_x) - load);
                load[x] += b->active[q].job->eng.weight + 1;
                xcc_of[q] = (uint32_t)x;
            }
            std::stable_sort(order[kd].begin(), order[kd].end(), [&](size_t x, size_t y) { return xcc_of[x] < xcc_of[y]; });
        }
        // A runtime failure from here on (the block, the launches, the read
        // back) leaves the launch's jobs undefined: some of their planned
        // phases may have run. They leave with the failure, each with its
        // own status, and the batch goes on with the others.
        auto round_ok = [&](hipError_t e, const char* what) {
            if (e == hipSuccess) return true;
            const int c = fail(e == hipErrorOutOfMemory ? TSG_ERR_OUT_OF_MEMORY : TSG_ERR_DEVICE,
                               std::string(what) + ": " + hipGetErrorString(e));
            for (size_t q = b->active.size(); q-- > 0;)
                if (std::find(runs.begin(), runs.end(), q) != runs.end()) eng_drop(b, q, c);
            return false;
        };
        // sizes of the launch block
        size_t n_jobs = runs.size(), n_ph = 0, n_args = 0, n_at = 0;
        for (size_t q : runs) {
            const tsg_job::Eng& E = b->active[q].job->eng;
            n_ph += E.prog.size();
            n_args += E.args.size();
            n_at += E.n_at.size();
        }
        const size_t o_ctl = 0, o_job = up256(2 * sizeof(QCtl)), o_dyn = up256(o_job + n_jobs * sizeof(QJob));
        const size_t o_ph = up256(o_dyn + n_jobs * sizeof(QDyn)), o_args = up256(o_ph + n_ph * sizeof(QPhase));
        const size_t o_nat = up256(o_args + n_args * sizeof(BatchArgs)), bytes = up256(o_nat + n_at * 4 + 4);
        if (bytes > b->h_cap) {
            // grown between launches (nothing reads the old block any more)
            const size_t cap = std::max(bytes, 2 * b->h_cap);
            char* h = nullptr;
            if (!round_ok(hipHostMalloc((void**)&h, cap), "hipHostMalloc(launch block)")) return TSG_OK;
            DBuf<char> d;
            const hipError_t e = d.alloc(cap);
            if (e != hipSuccess) {
                (void)hipHostFree(h);
                round_ok(e, "launch block on the device");
                return TSG_OK;
            }
            if (b->h_blk) (void)hipHostFree(b->h_blk);
            b->h_blk = h;
            b->h_cap = cap;
            std::swap(b->d_blk.p, d.p);
            std::swap(b->d_blk.n, d.n);
            std::swap(b->d_blk.bytes, d.bytes);
        }
        char* H = b->h_blk;
        char* D = b->d_blk.p;
        memset(H, 0, o_ph);
        QCtl* ctl = reinterpret_cast<QCtl*>(H + o_ctl);
        QJob* jobs = reinterpret_cast<QJob*>(H + o_job);
        QPhase* ph = reinterpret_cast<QPhase*>(H + o_ph);
        BatchArgs* args = reinterpret_cast<BatchArgs*>(H + o_args);
        uint32_t* nat = reinterpret_cast<uint32_t*>(H + o_nat);
        size_t jx = 0, px = 0, ax = 0, nx = 0;
        size_t first[2] = {0, 0}, count[2] = {0, 0};
        for (int kd = 0; kd < 2; ++kd) {
            first[kd] = jx;
            QCtl& C = ctl[kd];
            for (size_t q : order[kd]) {
                tsg_job* j = b->active[q].job;
                tsg_job::Eng& E = j->eng;
                const uint32_t x = xcc_of[q];
                QJob& J = jobs[jx];
                J.phases = reinterpret_cast<const QPhase*>(D + o_ph) + px;
                J.n_phases = (uint32_t)E.prog.size();
                J.xcc = x;
                J.rc = j->rc.p;
                J.result = j->result.p;
                J.overflow = j->overflow.p;
                J.mend = j->mend.p;
                J.tmin = j->tmin.p;
                J.n_at = reinterpret_cast<const uint32_t*>(D + o_nat) + nx;
                J.pre_nbrs = j->pre_nbrs.p;
                J.pre_cnt = j->pre_cnt.p;
                J.rank = j->rank.p;
                J.st_part = j->st_part.p;
                J.st_part_n = j->st_part_n.p;
                J.st_sum = j->st_sum.p;
                J.st_cnt = j->st_cnt.p;
                for (const QPhase& p0 : E.prog) {
                    QPhase p1 = p0;
                    p1.arg += (uint32_t)ax;
                    ph[px++] = p1;
                }
                memcpy(args + ax, E.args.data(), E.args.size() * sizeof(BatchArgs));
                ax += E.args.size();
                if (!E.n_at.empty()) memcpy(nat + nx, E.n_at.data(), E.n_at.size() * 4);
                nx += E.n_at.size();
                C.x[x].active += 1;
                for (uint32_t y = x + 1; y <= 8; ++y) C.job0[y] += 1;
                b->units_total += E.prog.size();
                ++jx;
            }
            count[kd] = jx - first[kd];
            for (uint32_t y = 0; y <= 8; ++y) C.job0[y] += (uint32_t)first[kd];
        }
        if (!round_ok(hipMemcpyAsync(D, H, o_nat + n_at * 4, hipMemcpyHostToDevice, s), "launch block upload"))
            return TSG_OK;
        lap(8);
        for (int kd = 0; kd < 2; ++kd) {
            if (!count[kd]) continue;
            if (!b->wgs[kd] && !round_ok(hipErrorInvalidValue, "engine occupancy query")) return TSG_OK;
            QEngineArgs ea{};
            ea.jobs = reinterpret_cast<const QJob*>(D + o_job);
            ea.dyn = reinterpret_cast<QDyn*>(D + o_dyn);
            ea.ctl = reinterpret_cast<QCtl*>(D + o_ctl) + kd;
            ea.offtab = b->ctx->offtab.p;
            ea.offtab_n = b->ctx->offtab_n;
            ea.idle_ticks = b->idle_ticks;
            ea.flags = spread[kd] ? kEngineSpread : 0u;
            const uint32_t wgs = b->wgs_cap ? std::min(b->wgs_cap, b->wgs[kd]) : b->wgs[kd];
            if (!round_ok(launch_engine(ea, reinterpret_cast<const BatchArgs*>(D + o_args), wgs, kd == 1, s),
                          "launch_engine"))
                return TSG_OK;
            b->launches += 1;
        }
        // control words and every job's cursor: each program must have run to its end
        std::vector<char> back(o_ph - o_ctl);
        if (!round_ok(hipMemcpyAsync(back.data(), D + o_ctl, back.size(), hipMemcpyDeviceToHost, s),
                      "control words read back") ||
            !round_ok(hipStreamSynchronize(s), "engine launch"))
            return TSG_OK;
        lap(9);
        const QCtl* bc = reinterpret_cast<const QCtl*>(back.data());
        uint32_t err[2] = {0, 0};
        for (int kd = 0; kd < 2; ++kd)
            for (int x = 0; x < 8; ++x) {
                const QXcd& X = bc[kd].x[x];
                err[kd] |= X.error;
                b->prof[0] += (double)X.t_find * 1e-5;
                b->prof[1] += (double)X.t_resolve * 1e-5;
                b->prof[2] += (double)X.t_other * 1e-5;
                b->prof[3] += (double)X.t_close * 1e-5;
                b->prof[4] += X.units_resolve;
                b->prof[5] += X.units_other;
            }
        const QDyn* bd = reinterpret_cast<const QDyn*>(back.data() + (o_dyn - o_ctl));
        for (size_t q = 0; q < n_jobs; ++q)
            if ((uint32_t)(bd[q].cursor >> 32) != jobs[q].n_phases) err[0] |= 2u;
        b->rounds += 1;
        for (size_t q : runs) {
            tsg_job::Eng& E = b->active[q].job->eng;
            E.prog.clear();
            E.args.clear();
            E.n_at.clear();
            E.weight = 0;
            E.items = 0;
        }
        if (err[0] | err[1]) {
            // a workgroup saw no progress on its XCD for the idle limit, or a
            // program did not run to its end: the launch's jobs are not trustworthy
            for (size_t q = b->active.size(); q-- > 0;)
                if (std::find(runs.begin(), runs.end(), q) != runs.end())
                    eng_drop(b, q, fail(TSG_ERR_DEVICE, "the work-queue engine stalled"));
            return TSG_OK;
        }
    }
    // jobs whose whole program has run
    for (size_t q = 0; q < b->active.size();) {
        tsg_job* j = b->active[q].job;
        if (j->eng.planned_all && j->eng.prog.empty()) {
            int rc = step_end(j);
            if (rc) {
                eng_drop(b, q, rc);
                continue;
            }
            b->done.push_back(b->active[q]);
            b->active.erase(b->active.begin() + q);
            continue;
        }
        ++q;
    }
    lap(10);
    return TSG_OK;
}

// Finalise jobs whose last launch has completed.
int batch_reap(tsg_batch* b, bool wait) {
    for (size_t q = 0; q < b->done.size();) {
        tsg_job* j = b->done[q].job;
        if (!wait && j->ev_run1 && hipEventQuery(j->ev_run1) == hipErrorNotReady) {
            ++q;
            continue;
        }
        if (b->done[q].alone) {
            // run_job has waited for its end and set its state
            j->step.active = false;
            j->ran = true;
        } else if (b->done[q].failed) {
            if (j->ev_run1) (void)hipEventSynchronize(j->ev_run1);
            j->step.active = false;
            j->ran = true;
        } else {
            b->done[q].status = step_finish(j);
        }
        b->taken.push_back(b->done[q]);
        b->done.erase(b->done.begin() + q);
    }
    return TSG_OK;
}

// Upload one image as RGBA8 packed u32.
std::vector<uint32_t> packed(const tsg_image& im) {
    std::vector<uint32_t> v((size_t)im.w * im.h);
    for (size_t i = 0; i < v.size(); ++i) v[i] = pack_px(im.rgba + i * 4);
    return v;
}

bool valid_image(const tsg_image* im) { return im && im->rgba && im->w > 0 && im->h > 0; }

}  // namespace

// ------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------
extern "C" {

void tsg_params_default(tsg_params* p) {
    p->nearest_neighbors = 50;
    p->random_sample_locations = 50;
    p->cauchy_dispersion = 1.0f;
    p->backtrack_percent = 0.5f;
    p->backtrack_stages = 5;
    p->guide_alpha = 0.8f;
    p->seed = 0;
    p->tiling_mode = 0;
    p->random_init = UINT64_MAX;
    p->max_thread_count = 1;
}

void tsg_schedule_default(tsg_schedule* s) {
    // per-stage mean cost within about 1.2% of one pixel per batch (seed-
    // averaged over 12 seeds at 256^2). The first stage's picks carry into
    // every later stage through the redo passes, so it gets the finer
    // new-pixel batches: (5, 6, 6) is 3-5% off in stages 2-5, (5, 7, 6) too.
    s->px_per_batch = 0;
    s->redo_shift = 5;
    s->new_shift = 6;
    s->max_batch = 0;
    s->first_new_shift = 7;
}

const char* tsg_last_error(void) { return g_err.c_str(); }
const char* tsg_version(void) { return "tsgpu 0.1 gfx950"; }

int tsg_ctx_create(int device, tsg_ctx** out) {
    if (!out) return fail(TSG_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(TSG_ERR_INVALID_ARGUMENT, "no such device");
    HIP_TRY(hipSetDevice(device));
    std::unique_ptr<tsg_ctx> c(new tsg_ctx());
    c->device = device;
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIP_TRY(c->staging.init(4u << 20, 4));
    // the second stream takes whole stages' pick orders, possibly several stages ahead
    HIP_TRY(c->staging_side.init(8u << 20, 8));
    HIP_TRY(configure_kernels());
    const std::vector<int32_t>& tab = offset_table();
    HIP_TRY(c->offtab.alloc(tab.size()));
    HIP_TRY(hipMemcpy(c->offtab.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice));
    c->offtab_n = (uint32_t)tab.size();
    HIP_TRY(c->dummy_lut.alloc(512));
    HIP_TRY(hipMemset(c->dummy_lut.p, 0, 512 * 4));
    void* sp = nullptr;
    HIP_TRY(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
    c->scratch = static_cast<unsigned char*>(sp);
    *out = c.release();
    return TSG_OK;
}

void tsg_ctx_destroy(tsg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    ctx->offtab.release();
    ctx->dummy_lut.release();
    // the device's cached blocks are idle: give them back
    if (DevPool* pl = pool_here()) {
        std::lock_guard<std::mutex> lk(pl->mu);
        for (auto& kv : pl->free) (void)hipFree(kv.second);
        pl->free.clear();
        pl->cached = 0;
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    delete ctx;
}

int tsg_job_create(tsg_ctx* ctx, const tsg_params* prm, uint32_t n_examples, const tsg_image* examples,
                   const int* sample_kind, const tsg_image* sample_masks, const tsg_image* example_guides,
                   const tsg_image* target_guide, const tsg_image* inpaint_mask, uint32_t inpaint_index,
                   uint32_t out_w, uint32_t out_h, tsg_job** out) {
    if (!ctx || !prm || !out) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    // check_parameters_validity (session.rs:450-499)
    int rc;
    if ((rc = check_range(prm->cauchy_dispersion, 0.0f, 1.0f, "cauchy-dispersion"))) return rc;
    if ((rc = check_range(prm->backtrack_percent, 0.0f, 1.0f, "backtrack-percent"))) return rc;
    if ((rc = check_range(prm->guide_alpha, 0.0f, 1.0f, "guide-alpha"))) return rc;
    if (prm->max_thread_count == 0) return check_range(0.0f, 1.0f, 1024.0f, "max-thread-count");
    if (prm->random_sample_locations == 0) return check_range(0.0f, 1.0f, 1024.0f, "m-rand");
    // check_images_validity (session.rs:501-524)
    if (n_examples == 0 || !examples) return fail(TSG_ERR_NO_EXAMPLES, "at least 1 example must be available as a sampling source");
    uint32_t usable = 0;
    for (uint32_t i = 0; i < n_examples; ++i)
        if (!sample_kind || sample_kind[i] != TSG_SAMPLE_IGNORE) ++usable;
    if (usable == 0) return fail(TSG_ERR_NO_EXAMPLES, "at least 1 example must be available as a sampling source");
    if (example_guides && !target_guide)
        return fail(TSG_ERR_UNSUPPORTED, "example guides without a target guide are not used by the reference");
    // device-path limits
    if (prm->nearest_neighbors == 0 || prm->nearest_neighbors > (uint32_t)kMaxK)
        return fail(TSG_ERR_UNSUPPORTED, "nearest_neighbors must be in 1..64 on the device");
    if (prm->random_sample_locations > 4096) return fail(TSG_ERR_UNSUPPORTED, "random_sample_locations above 4096");
    if (out_w == 0 || out_h == 0 || out_w > 16384 || out_h > 16384)
        return fail(TSG_ERR_INVALID_ARGUMENT, "output size must be 1..16384 per side");
    for (uint32_t i = 0; i < n_examples; ++i) {
        if (!valid_image(&examples[i])) return fail(TSG_ERR_INVALID_ARGUMENT, "invalid example image");
        if (examples[i].w > 65535 || examples[i].h > 65535)
            return fail(TSG_ERR_UNSUPPORTED, "example side above 65535");
        // the gathers address an example level by a 32-bit byte offset
        if ((uint64_t)examples[i].w * examples[i].h > (1ull << 30))
            return fail(TSG_ERR_UNSUPPORTED, "example above 2^30 pixels");
        int kd = sample_kind ? sample_kind[i] : TSG_SAMPLE_ALL;
        if (kd == TSG_SAMPLE_IMAGE) {
            if (!sample_masks || !valid_image(&sample_masks[i]))
                return fail(TSG_ERR_INVALID_ARGUMENT, "missing sample mask");
            if (sample_masks[i].w < examples[i].w || sample_masks[i].h < examples[i].h)
                return fail(TSG_ERR_SIZE_MISMATCH, "sample mask smaller than its example");
        }
        if (example_guides && !valid_image(&example_guides[i]))
            return fail(TSG_ERR_INVALID_ARGUMENT, "invalid example guide");
        if (example_guides && (uint64_t)example_guides[i].w * example_guides[i].h > (1ull << 30))
            return fail(TSG_ERR_UNSUPPORTED, "example guide above 2^30 pixels");
    }
    if (target_guide && (!valid_image(target_guide) || target_guide->w != out_w || target_guide->h != out_h))
        return fail(TSG_ERR_SIZE_MISMATCH, "target guide must have the output size");
    if (inpaint_mask) {
        if (!valid_image(inpaint_mask) || inpaint_mask->w != out_w || inpaint_mask->h != out_h)
            return fail(TSG_ERR_SIZE_MISMATCH, "inpaint mask must have the output size");
        if (inpaint_index >= n_examples) return fail(TSG_ERR_INVALID_ARGUMENT, "inpaint index out of range");
        if (examples[inpaint_index].w != out_w || examples[inpaint_index].h != out_h)
            return fail(TSG_ERR_SIZE_MISMATCH, "the inpainted example must have the output size");
    }
    if (prm->random_init != UINT64_MAX && usable != n_examples)
        return fail(TSG_ERR_UNSUPPORTED, "random_init with ignored examples mixes map numbering (reference quirk)");

    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<tsg_job> j(new tsg_job());
    j->ctx = ctx;
    j->prm = *prm;
    tsg_schedule_default(&j->sched);
    tsg_tuning_default(&j->tune);
    j->W = out_w;
    j->H = out_h;
    j->N = (size_t)out_w * out_h;
    j->tiles_x = (out_w + 7) / 8;
    j->tiles_y = (out_h + 7) / 8;
    j->n_ex = n_examples;
    j->levels = prm->backtrack_stages == 0 ? 1 : prm->backtrack_stages;
    j->guided = target_guide != nullptr;
    for (uint32_t i = 0; i < n_examples; ++i) {
        int kd = sample_kind ? sample_kind[i] : TSG_SAMPLE_ALL;
        j->kind.push_back(kd);
        if (kd != TSG_SAMPLE_IGNORE) j->filt.push_back(i);
    }
    if (inpaint_mask && inpaint_index >= j->filt.size())
        return fail(TSG_ERR_UNSUPPORTED, "inpaint example index beyond the sampled examples (the reference panics)");
    hipStream_t s = ctx->stream;
    std::lock_guard<std::mutex> lk(ctx->mu);
    // style transfer (session.rs:385-393, lib.rs:572-581): without example
    // guides the target guide becomes a guide map, and each example's guide
    // is its own guide map matched to the target's histogram
    std::vector<std::vector<uint8_t>> sg_ex;
    std::vector<uint8_t> sg_tg;
    std::vector<tsg_image> sg_imgs;
    tsg_image sg_target{};
    if (target_guide && !example_guides) {
        if ((rc = style_guide(s, *target_guide, sg_tg))) return rc;
        sg_target = tsg_image{target_guide->w, target_guide->h, sg_tg.data()};
        sg_ex.resize(n_examples);
        sg_imgs.resize(n_examples);
        for (uint32_t i = 0; i < n_examples; ++i) {
            if ((rc = style_guide(s, examples[i], sg_ex[i]))) return rc;
            match_histograms(sg_ex[i], sg_tg);
            sg_imgs[i] = tsg_image{examples[i].w, examples[i].h, sg_ex[i].data()};
        }
        example_guides = sg_imgs.data();
        target_guide = &sg_target;
    }
    for (uint32_t i = 0; i < n_examples; ++i) {
        const tsg_image& im = examples[i];
        j->ex_dims.push_back({im.w, im.h});
        j->h_finest.push_back(packed(im));
        std::unique_ptr<DBuf<uint32_t>> b(new DBuf<uint32_t>());
        HIP_TRY(b->alloc((size_t)j->levels * im.w * im.h));
        if ((rc = device_pyramid(im.rgba, im.w, im.h, j->levels, b->p, s))) return rc;
        j->ex.push_back(std::move(b));
        std::unique_ptr<DBuf<uint8_t>> mk(new DBuf<uint8_t>());
        if (j->kind[i] == TSG_SAMPLE_IMAGE) {
            const tsg_image& sm = sample_masks[i];
            std::vector<uint8_t> v((size_t)im.w * im.h);
            size_t valid = 0;
            for (uint32_t y = 0; y < im.h; ++y)
                for (uint32_t x = 0; x < im.w; ++x) {
                    uint8_t r = sm.rgba[((size_t)y * sm.w + x) * 4];
                    v[(size_t)y * im.w + x] = r != 0;
                    valid += r != 0;
                }
            if (valid == 0)
                return fail(TSG_ERR_SAMPLING, "sample mask admits no pixel (the reference loops forever)");
            HIP_TRY(mk->alloc(v.size()));
            HIP_TRY(hipMemcpy(mk->p, v.data(), v.size(), hipMemcpyHostToDevice));
        }
        j->smask.push_back(std::move(mk));
        if (example_guides) {
            const tsg_image& g = example_guides[i];
            j->g_dims.push_back({g.w, g.h});
            std::unique_ptr<DBuf<uint32_t>> gb(new DBuf<uint32_t>());
            HIP_TRY(gb->alloc((size_t)j->levels * g.w * g.h));
            if ((rc = device_pyramid(g.rgba, g.w, g.h, j->levels, gb->p, s))) return rc;
            j->exg.push_back(std::move(gb));
        }
    }
    if (target_guide) {
        HIP_TRY(j->tg.alloc((size_t)j->levels * j->N));
        if ((rc = device_pyramid(target_guide->rgba, out_w, out_h, j->levels, j->tg.p, s))) return rc;
    }
    // map descriptors per level; the guide of filtered map m is the m-th
    // entry of the unfiltered guide list (get_single_guide_level, ms.rs:1566-1573)
    const uint32_t nf = (uint32_t)j->filt.size();
    std::vector<MapDesc> md((size_t)j->levels * nf);
    for (uint32_t lv = 0; lv < j->levels; ++lv)
        for (uint32_t mi = 0; mi < nf; ++mi) {
            uint32_t e = j->filt[mi];
            MapDesc d{};
            d.w = j->ex_dims[e].first;
            d.h = j->ex_dims[e].second;
            d.px = j->ex[e]->p + (size_t)lv * d.w * d.h;
            d.smask = j->smask[e]->p;
            if (j->guided) {
                d.gw = j->g_dims[mi].first;
                d.gh = j->g_dims[mi].second;
                d.gpx = j->exg[mi]->p + (size_t)lv * d.gw * d.gh;
            }
            md[(size_t)lv * nf + mi] = d;
        }
    HIP_TRY(j->maps.alloc(md.size()));
    HIP_TRY(hipMemcpy(j->maps.p, md.data(), md.size() * sizeof(MapDesc), hipMemcpyHostToDevice));
    if (inpaint_mask) {
        j->inpaint = true;
        j->inpaint_index = inpaint_index;
        j->inpaint_mask_r.resize(j->N);
        for (size_t i = 0; i < j->N; ++i) j->inpaint_mask_r[i] = inpaint_mask->rgba[i * 4];
    }
    const uint32_t P = prm->backtrack_stages;
    HIP_TRY(j->state.alloc(j->N));
    HIP_TRY(j->mask.alloc((size_t)j->tiles_x * j->tiles_y));
    if (prm->tiling_mode) HIP_TRY(j->gmask.alloc((size_t)j->tiles_x * j->tiles_y));
    HIP_TRY(j->score.alloc(j->N));
    HIP_TRY(j->rlist.alloc(j->N));
    HIP_TRY(j->item_score.alloc(j->N));
    HIP_TRY(j->lut.alloc(512));
    HIP_TRY(j->overflow.alloc(1));
    HIP_TRY(j->st_sum.alloc(P + 1));
    HIP_TRY(j->st_cnt.alloc(P + 1));
    HIP_TRY(j->st_part.alloc((j->N + kStatsChunk - 1) / kStatsChunk));
    HIP_TRY(j->st_part_n.alloc((j->N + kStatsChunk - 1) / kStatsChunk));
    HIP_TRY(j->done_counter.alloc(1));
    HIP_TRY(j->rank.alloc(j->N));
    HIP_TRY(j->mend.alloc((size_t)j->tiles_x * j->tiles_y));
    HIP_TRY(j->tmin.alloc((size_t)j->tiles_x * j->tiles_y));
    HIP_TRY(hipMemset(j->done_counter.p, 0, 4));
    // Generator::new / new_from_inpaint and the random init are content-
    // independent of the stages: upload the initial state now
    if ((rc = init_state(j.get()))) return rc;
    j->state_ready = true;
    HIP_TRY(hipStreamSynchronize(s));
    {
        // a launch structure the autotuner stored for this shape (ctx->mu is held)
        auto it = ctx->tunings.find(tuning_key(j.get()));
        if (it != ctx->tunings.end()) j->tune = it->second;
    }
    *out = j.release();
    return TSG_OK;
}

void tsg_tuning_default(tsg_tuning* t) {
    // measured on the 512^2 -> 1024^2 workload: the last-workgroup commit
    // (a device-wide fence and counter per workgroup) costs more than a
    // separate commit launch; the one-workgroup loop pays only while batches
    // leave most of its 16 waves to teams
    t->small_max = 4;
    t->fuse_max = 0;
    t->team_max = 8;
    t->team_waves = 4096;
    t->prefetch_max = 1024;  // autotuned: 26.3 ms against 27.2 ms at 2048
    // 2 (a set per stage, every neighbour list ahead) measured 27.2 ms against
    // 26.9 ms: the second stream's kernels wait for free slots behind the
    // latency-bound launches instead of filling them
    t->ahead = 1;
    t->coop_wgs = 0;  // a grid barrier costs more than the launches it saves (measured)
    t->coop_max = 0;
    t->flow = 1;  // 38.9 -> 27.9 ms on that workload
    t->flow_wgs = 0;
    t->flow_team = 4;
    // 1 measured 26.99 ms against 26.34 ms (1024^2) and 71.5 against 67.8 ms
    // (2048^2), 3% faster at 256^2 and 512^2: it saves the commit dispatch
    // (5.4 us) but each resolve dispatch grows by 7.6 us
    t->defer = 0;
    // 0: with three 8-wave workgroups resident per CU (unguided), the
    // per-8-items launch measured 24.05 ms against 25.1 ms for the persistent
    // one, which its 128 VGPRs hold at two workgroups
    t->persist = 0;
    // 25.6 ms against 26.1 ms (1024-item heads of 256..2048 measured alike;
    // 4-wave heads 25.9, 2-wave heads 26.7)
    t->flow_head = 1024;
    t->flow_head_team = 8;
}

int tsg_job_set_tuning(tsg_job* job, const tsg_tuning* t) {
    if (!job || !t) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_tuning(t);
    if (rc) return rc;
    job->tune = *t;
    return TSG_OK;
}

int tsg_job_get_tuning(const tsg_job* job, tsg_tuning* out) {
    if (!job || !out) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    *out = job->tune;
    return TSG_OK;
}

int tsg_ctx_store_tuning(tsg_ctx* ctx, const tsg_job* like, const tsg_tuning* t) {
    if (!ctx || !like || !t) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    if (like->ctx != ctx) return fail(TSG_ERR_INVALID_ARGUMENT, "the job belongs to another context");
    int rc = check_tuning(t);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->tunings[tuning_key(like)] = *t;
    return TSG_OK;
}

int tsg_ctx_clear_tunings(tsg_ctx* ctx) {
    if (!ctx) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->tunings.clear();
    return TSG_OK;
}

int tsg_job_set_schedule(tsg_job* job, const tsg_schedule* s) {
    if (!job || !s) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    if (s->redo_shift > 31 || s->new_shift > 31 || s->first_new_shift > 31)
        return fail(TSG_ERR_INVALID_ARGUMENT, "shift above 31");
    job->sched = *s;
    return TSG_OK;
}

int tsg_job_run(tsg_job* job, tsg_progress_fn progress, void* user) {
    if (!job) return fail(TSG_ERR_INVALID_ARGUMENT, "null job");
    HIP_TRY(hipSetDevice(job->ctx->device));
    std::lock_guard<std::mutex> lk(job->ctx->mu);
    return run_job(job, progress, user);
}

int tsg_job_download_packed(tsg_job* job, uint32_t* state, uint32_t* resolved_flat, float* resolved_score) {
    if (!job) return fail(TSG_ERR_INVALID_ARGUMENT, "null job");
    HIP_TRY(hipSetDevice(job->ctx->device));
    std::lock_guard<std::mutex> lk(job->ctx->mu);
    if (job->step.active) return fail(TSG_ERR_INVALID_ARGUMENT, "the job is still running");
    // the job's work has completed: the download uses the side stream and so
    // overlaps whatever else the job stream is running (other jobs' rounds)
    hipStream_t s = job->ctx->side;
    const size_t N = job->N, R = job->n_res;
    float* rs = job->item_score.p;  // per-item scratch, free once the job has run
    if (resolved_score && R) {
        UnpackArgs ua{job->state.p, 0u, nullptr, nullptr, nullptr, job->rlist.p, job->score.p, (uint32_t)R, rs};
        HIP_TRY(launch_unpack(ua, s));
    }
    // device -> pinned landing area -> caller's buffers
    const size_t b_state = state ? N * sizeof(uint4) : 0, b_flat = resolved_flat ? R * 4 : 0;
    const size_t b_score = resolved_score ? R * 4 : 0;
    Landing& L = job->ctx->landing;
    HIP_TRY(L.reserve(b_state + b_flat + b_score + 1));
    if (b_state) HIP_TRY(hipMemcpyAsync(L.buf, job->state.p, b_state, hipMemcpyDeviceToHost, s));
    if (b_flat) HIP_TRY(hipMemcpyAsync(L.buf + b_state, job->rlist.p, b_flat, hipMemcpyDeviceToHost, s));
    if (b_score) HIP_TRY(hipMemcpyAsync(L.buf + b_state + b_flat, rs, b_score, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const unsigned nt = std::max(1u, std::min(8u, std::thread::hardware_concurrency() / 2));
    if (b_state) parallel_copy(state, L.buf, b_state, nt);
    if (b_flat) parallel_copy(resolved_flat, L.buf + b_state, b_flat, nt);
    if (b_score) parallel_copy(resolved_score, L.buf + b_state + b_flat, b_score, nt);
    return TSG_OK;
}

int tsg_job_debug_maps(tsg_job* job, uint8_t* patch_ids, uint8_t* map_ids, uint8_t* uncertainty) {
    if (!job) return fail(TSG_ERR_INVALID_ARGUMENT, "null job");
    HIP_TRY(hipSetDevice(job->ctx->device));
    std::lock_guard<std::mutex> lk(job->ctx->mu);
    if (job->step.active || !job->state_ready) return fail(TSG_ERR_INVALID_ARGUMENT, "the job has no finished state");
    hipStream_t s = job->ctx->side;
    const size_t N = job->N;
    DBuf<uint32_t> out;
    HIP_TRY(out.alloc(3 * N));
    uint32_t* dp = patch_ids ? out.p : nullptr;
    uint32_t* dm = map_ids ? out.p + N : nullptr;
    HIP_TRY(launch_id_maps(job->state.p, (uint32_t)N, dp, dm, s));
    if (uncertainty) {
        HIP_TRY(hipMemsetAsync(out.p + 2 * N, 0, N * 4, s));
        HIP_TRY(launch_uncertainty(job->rlist.p, job->score.p, (uint32_t)job->n_res, out.p + 2 * N, s));
    }
    if (patch_ids) HIP_TRY(hipMemcpyAsync(patch_ids, out.p, N * 4, hipMemcpyDeviceToHost, s));
    if (map_ids) HIP_TRY(hipMemcpyAsync(map_ids, out.p + N, N * 4, hipMemcpyDeviceToHost, s));
    if (uncertainty) HIP_TRY(hipMemcpyAsync(uncertainty, out.p + 2 * N, N * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return TSG_OK;
}

int tsg_xform_apply(tsg_ctx* ctx, const uint32_t* coord, uint32_t w, uint32_t h, uint32_t n_sources,
                    const tsg_image* sources, uint8_t* out) {
    if (!ctx || !coord || !out || (n_sources && !sources)) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    if (w == 0 || h == 0) return fail(TSG_ERR_INVALID_ARGUMENT, "empty output");
    for (uint32_t m = 0; m < n_sources; ++m)
        if (!valid_image(&sources[m])) return fail(TSG_ERR_INVALID_ARGUMENT, "invalid source image");
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->side;
    const size_t n = (size_t)w * h;
    size_t px = 0;
    for (uint32_t m = 0; m < n_sources; ++m) px += (size_t)sources[m].w * sources[m].h;
    DBuf<uint32_t> dcoord, dpx, dout, bad;
    DBuf<XformSource> dsrc;
    HIP_TRY(dcoord.alloc(3 * n));
    HIP_TRY(dpx.alloc(px));
    HIP_TRY(dout.alloc(n));
    HIP_TRY(bad.alloc(1));
    HIP_TRY(dsrc.alloc(std::max(1u, n_sources)));
    std::vector<XformSource> hs(n_sources);
    size_t off = 0;
    for (uint32_t m = 0; m < n_sources; ++m) {
        const size_t np = (size_t)sources[m].w * sources[m].h;
        // RGBA8 rows are the packed u32 layout of the device maps
        HIP_TRY(hipMemcpyAsync(dpx.p + off, sources[m].rgba, np * 4, hipMemcpyHostToDevice, s));
        hs[m] = XformSource{dpx.p + off, sources[m].w, sources[m].h};
        off += np;
    }
    if (n_sources) HIP_TRY(hipMemcpyAsync(dsrc.p, hs.data(), n_sources * sizeof(XformSource), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dcoord.p, coord, 3 * n * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(bad.p, 0, 4, s));
    HIP_TRY(launch_xform_apply(dcoord.p, (uint32_t)n, dsrc.p, n_sources, dout.p, bad.p, s));
    uint32_t hb = 0;
    HIP_TRY(hipMemcpyAsync(out, dout.p, n * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&hb, bad.p, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (hb) return fail(TSG_ERR_INVALID_ARGUMENT, "a coordinate lies outside its source map (the reference panics)");
    return TSG_OK;
}

int tsg_job_download(tsg_job* job, uint8_t* rgba, uint32_t* coord, uint32_t* ids, uint32_t* resolved_flat,
                     float* resolved_score) {
    if (!job) return fail(TSG_ERR_INVALID_ARGUMENT, "null job");
    HIP_TRY(hipSetDevice(job->ctx->device));
    std::lock_guard<std::mutex> lk(job->ctx->mu);
    if (job->step.active) return fail(TSG_ERR_INVALID_ARGUMENT, "the job is still running");
    hipStream_t s = job->ctx->side;  // the job's work has completed (see tsg_job_download_packed)
    // unpack on the device (8 + 12 + 8 + 4 bytes per pixel), then one copy per map
    const size_t N = job->N, R = job->n_res;
    DBuf<uint32_t> out;
    HIP_TRY(out.alloc(6 * N + R));
    UnpackArgs ua{job->state.p, (uint32_t)N, rgba ? out.p : nullptr, coord ? out.p + N : nullptr,
                  ids ? out.p + 4 * N : nullptr, job->rlist.p, job->score.p, resolved_score ? (uint32_t)R : 0u,
                  resolved_score ? reinterpret_cast<float*>(out.p + 6 * N) : nullptr};
    HIP_TRY(launch_unpack(ua, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (rgba) HIP_TRY(hipMemcpy(rgba, out.p, N * 4, hipMemcpyDeviceToHost));
    if (coord) HIP_TRY(hipMemcpy(coord, out.p + N, N * 12, hipMemcpyDeviceToHost));
    if (ids) HIP_TRY(hipMemcpy(ids, out.p + 4 * N, N * 8, hipMemcpyDeviceToHost));
    if (resolved_flat && R) HIP_TRY(hipMemcpy(resolved_flat, job->rlist.p, R * 4, hipMemcpyDeviceToHost));
    if (resolved_score && R) HIP_TRY(hipMemcpy(resolved_score, out.p + 6 * N, R * 4, hipMemcpyDeviceToHost));
    return TSG_OK;
}

int tsg_job_stage_stats(tsg_job* job, double* sums, uint64_t* counts, uint64_t* items, uint32_t cap) {
    if (!job) return fail(TSG_ERR_INVALID_ARGUMENT, "null job");
    uint32_t n = std::min<uint32_t>(cap, job->prm.backtrack_stages + 1);
    std::vector<double> s(job->prm.backtrack_stages + 1);
    std::vector<unsigned long long> c(job->prm.backtrack_stages + 1);
    HIP_TRY(hipSetDevice(job->ctx->device));
    HIP_TRY(hipMemcpy(s.data(), job->st_sum.p, s.size() * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(c.data(), job->st_cnt.p, c.size() * 8, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) {
        if (sums) sums[i] = s[i];
        if (counts) counts[i] = c[i];
        if (items) items[i] = i < job->st_items.size() ? job->st_items[i] : 0;
    }
    return (int)n;
}

int tsg_job_timing(tsg_job* job, double* run_ms, uint64_t* n_batches) {
    if (!job) return fail(TSG_ERR_INVALID_ARGUMENT, "null job");
    if (run_ms) *run_ms = job->run_ms;
    if (n_batches) *n_batches = job->n_batches;
    return TSG_OK;
}

int tsg_job_set_profiling(tsg_job* job, int level) {
    if (!job) return fail(TSG_ERR_INVALID_ARGUMENT, "null job");
    job->profile = level != 0;
    job->profile_level = level;
    if (level >= 2) {
        HIP_TRY(hipSetDevice(job->ctx->device));
        HIP_TRY(job->clk.reserve(job->N * 16));
        HIP_TRY(hipMemset(job->clk.p, 0, job->N * 16 * 8));
    }
    return TSG_OK;
}

int tsg_job_phase_clocks(tsg_job* job, uint64_t* out, uint64_t n_items) {
    if (!job || !out) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    if (job->profile_level < 2) return fail(TSG_ERR_INVALID_ARGUMENT, "phase clocks need profiling level 2");
    n_items = std::min<uint64_t>(n_items, job->N);
    HIP_TRY(hipMemcpy(out, job->clk.p, n_items * 16 * 8, hipMemcpyDeviceToHost));
    return TSG_OK;
}

int tsg_job_kernel_stats(tsg_job* job, double* ms, uint64_t* items, uint64_t* launches) {
    if (!job) return fail(TSG_ERR_INVALID_ARGUMENT, "null job");
    for (int q = 0; q < kKinds; ++q) {
        if (ms) ms[q] = job->kern_ms[q];
        if (items) items[q] = job->kern_items[q];
        if (launches) launches[q] = job->kern_launches[q];
    }
    return TSG_OK;
}

void tsg_job_destroy(tsg_job* job) {
    if (!job) return;
    if (job->owner) {
        // still held by a batch, which destroys it (or hands it back with tsg_batch_take)
        fail(TSG_ERR_INVALID_ARGUMENT, "the job belongs to a batch until it is taken back");
        return;
    }
    (void)hipSetDevice(job->ctx->device);
    // a finished job has no work in flight (downloads wait for theirs)
    if (job->step.active) {
        (void)hipStreamSynchronize(job->ctx->stream);
        (void)hipStreamSynchronize(job->ctx->side);
    }
    // its blocks go to the device's cache
    g_pool_release = true;
    delete job;
    g_pool_release = false;
}

int tsg_pyramid_build(tsg_ctx* ctx, const tsg_image* img, uint32_t levels, uint8_t* out) {
    if (!ctx || !valid_image(img) || !out) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lk(ctx->mu);
    uint32_t L = levels == 0 ? 1 : levels;
    size_t npx = (size_t)img->w * img->h;
    DBuf<uint32_t> d;
    HIP_TRY(d.alloc(L * npx));
    int rc = device_pyramid(img->rgba, img->w, img->h, L, d.p, ctx->stream);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy(out, d.p, L * npx * 4, hipMemcpyDeviceToHost));
    return TSG_OK;
}

int tsg_resize(tsg_ctx* ctx, const tsg_image* img, uint32_t nw, uint32_t nh, int filter, uint8_t* out) {
    if (!ctx || !valid_image(img) || !out || nw == 0 || nh == 0)
        return fail(TSG_ERR_INVALID_ARGUMENT, "null argument or zero size");
    if (filter < TSG_FILTER_TRIANGLE || filter > TSG_FILTER_GAUSSIAN)
        return fail(TSG_ERR_INVALID_ARGUMENT, "unknown resampling filter");
    const size_t n_in = (size_t)img->w * img->h, n_out = (size_t)nw * nh;
    if (img->w == nw && img->h == nh) {  // load_image keeps an image of the asked size as it is (utils.rs:66)
        std::memcpy(out, img->rgba, n_in * 4);
        return TSG_OK;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lk(ctx->mu);
    const Filter f = filter == TSG_FILTER_TRIANGLE ? Filter{kern_triangle, 1.0f}
                     : filter == TSG_FILTER_CATMULLROM ? Filter{kern_catmullrom, 2.0f}
                                                       : Filter{kern_gaussian, 3.0f};
    DBuf<uint32_t> src, dst;
    HIP_TRY(src.alloc(n_in));
    HIP_TRY(dst.alloc(n_out));
    HIP_TRY(hipMemcpyAsync(src.p, img->rgba, n_in * 4, hipMemcpyHostToDevice, ctx->stream));
    int rc = device_resize(src.p, img->w, img->h, dst.p, nw, nh, f, ctx->stream);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(out, dst.p, n_out * 4, hipMemcpyDeviceToHost));
    return TSG_OK;
}

int tsg_knn_batch(tsg_ctx* ctx, const uint8_t* mask, uint32_t w, uint32_t h, uint32_t n_q, const uint32_t* q_flat,
                  uint32_t k, int32_t* nbrs, uint32_t* counts) {
    if (!ctx || !mask || !q_flat || !nbrs || !counts) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    if (k == 0 || k > (uint32_t)kMaxK) return fail(TSG_ERR_UNSUPPORTED, "k must be in 1..64");
    if (w == 0 || h == 0 || w > 16384 || h > 16384) return fail(TSG_ERR_INVALID_ARGUMENT, "bad size");
    for (uint32_t q = 0; q < n_q; ++q)
        if (q_flat[q] >= w * h) return fail(TSG_ERR_INVALID_ARGUMENT, "query outside the image");
    if (n_q == 0) return TSG_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    uint32_t tx = (w + 7) / 8, ty = (h + 7) / 8;
    std::vector<uint64_t> tiles((size_t)tx * ty, 0);
    std::vector<uint32_t> rl;
    for (uint32_t i = 0; i < w * h; ++i)
        if (mask[i]) {
            uint32_t x = i % w, y = i / w;
            tiles[(y >> 3) * tx + (x >> 3)] |= 1ull << (((y & 7u) << 3) | (x & 7u));
            rl.push_back(i);
        }
    DBuf<uint64_t> dm;
    DBuf<uint32_t> drl, dq, dcnt;
    DBuf<int32_t> dn;
    DBuf<uint4> dres;
    HIP_TRY(dm.alloc(tiles.size()));
    HIP_TRY(drl.alloc(rl.size()));
    HIP_TRY(dq.alloc(n_q));
    HIP_TRY(dcnt.alloc(n_q));
    HIP_TRY(dn.alloc((size_t)n_q * k));
    HIP_TRY(dres.alloc(n_q));
    HIP_TRY(hipMemcpy(dm.p, tiles.data(), tiles.size() * 8, hipMemcpyHostToDevice));
    if (!rl.empty()) HIP_TRY(hipMemcpy(drl.p, rl.data(), rl.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dq.p, q_flat, n_q * 4, hipMemcpyHostToDevice));
    BatchArgs a{};
    a.W = w;
    a.H = h;
    a.tiles_x = tx;
    a.tiles_y = ty;
    a.mask = dm.p;
    a.work = dq.p;
    a.rlist = drl.p;
    a.n_res = (uint32_t)rl.size();
    a.offtab = ctx->offtab.p;
    a.offtab_n = ctx->offtab_n;
    a.lut = ctx->dummy_lut.p;
    a.item0 = 0;
    a.item1 = n_q;
    a.k = k;
    a.result = dres.p;
    a.dbg_nbrs = dn.p;
    a.dbg_nbr_count = dcnt.p;
    a.knn_only = 1;
    HIP_TRY(launch_resolve(a, false, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<int32_t> packedn((size_t)n_q * k);
    HIP_TRY(hipMemcpy(packedn.data(), dn.p, packedn.size() * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(counts, dcnt.p, n_q * 4, hipMemcpyDeviceToHost));
    for (uint32_t q = 0; q < n_q; ++q) {
        int x = (int)(q_flat[q] % w), y = (int)(q_flat[q] / w);
        for (uint32_t jn = 0; jn < k; ++jn) {
            int32_t o = packedn[(size_t)q * k + jn];
            bool v = jn < counts[q];
            nbrs[((size_t)q * k + jn) * 2] = v ? x + (int)(int16_t)(o & 0xffff) : 0;
            nbrs[((size_t)q * k + jn) * 2 + 1] = v ? y + (o >> 16) : 0;
        }
    }
    return TSG_OK;
}

int tsg_knn_batch_noalloc(tsg_ctx* ctx, const uint8_t* mask, uint32_t w, uint32_t h, uint32_t n_q,
                          const uint32_t* q_flat, uint32_t k, int32_t* nbrs, uint32_t* counts) {
    if (!ctx || !mask || !q_flat || !nbrs || !counts) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    if (k == 0 || k > (uint32_t)kMaxK) return fail(TSG_ERR_UNSUPPORTED, "k must be in 1..64");
    if (w == 0 || h == 0 || w > 16384 || h > 16384) return fail(TSG_ERR_INVALID_ARGUMENT, "bad size");
    for (uint32_t q = 0; q < n_q; ++q)
        if (q_flat[q] >= w * h) return fail(TSG_ERR_INVALID_ARGUMENT, "query outside the image");
    if (n_q == 0) return TSG_OK;
    // scratch layout: mask bytes | tile bits | resolved list | queries | counts | packed lists | results | pairs
    const uint32_t tx = (w + 7) / 8, ty = (h + 7) / 8;
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t o_mask = 0;
    const size_t o_tiles = up(o_mask + (size_t)w * h);
    const size_t o_rl = up(o_tiles + (size_t)tx * ty * 8);
    const size_t o_q = up(o_rl + (size_t)kBruteMax * 4);
    const size_t o_cnt = up(o_q + (size_t)n_q * 4);
    const size_t o_packed = up(o_cnt + (size_t)n_q * 4);
    const size_t o_res = up(o_packed + (size_t)n_q * k * 4);
    const size_t o_pairs = up(o_res + (size_t)n_q * sizeof(uint4));
    const size_t total = o_pairs + (size_t)n_q * k * 8;
    if (total > kScratchBytes)
        return fail(TSG_ERR_UNSUPPORTED, "mask and queries need " + std::to_string(total) +
                                             " bytes of device scratch; the no-allocation entry point has " +
                                             std::to_string(kScratchBytes) + " (use tsg_knn_batch)");
    if (!ctx->scratch) return fail(TSG_ERR_DEVICE, "device scratch address unavailable");
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    // the resolved list is read only while it holds at most kBruteMax points
    uint32_t rl[kBruteMax];
    uint32_t n_res = 0;
    for (uint32_t i = 0; i < w * h; ++i)
        if (mask[i]) {
            if (n_res < kBruteMax) rl[n_res] = i;
            ++n_res;
        }
    unsigned char* base = ctx->scratch;
    HIP_TRY(hipMemcpyAsync(base + o_mask, mask, (size_t)w * h, hipMemcpyHostToDevice, s));
    HIP_TRY(launch_mask_tiles(base + o_mask, w, h, tx, ty, reinterpret_cast<uint64_t*>(base + o_tiles), s));
    if (n_res > 0 && n_res <= kBruteMax)
        HIP_TRY(hipMemcpyAsync(base + o_rl, rl, (size_t)n_res * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(base + o_q, q_flat, (size_t)n_q * 4, hipMemcpyHostToDevice, s));
    BatchArgs a{};
    a.W = w;
    a.H = h;
    a.tiles_x = tx;
    a.tiles_y = ty;
    a.mask = reinterpret_cast<const uint64_t*>(base + o_tiles);
    a.work = reinterpret_cast<const uint32_t*>(base + o_q);
    a.rlist = reinterpret_cast<const uint32_t*>(base + o_rl);
    a.n_res = n_res;
    a.offtab = ctx->offtab.p;
    a.offtab_n = ctx->offtab_n;
    a.lut = ctx->dummy_lut.p;
    a.item0 = 0;
    a.item1 = n_q;
    a.k = k;
    a.result = reinterpret_cast<uint4*>(base + o_res);
    a.dbg_nbrs = reinterpret_cast<int32_t*>(base + o_packed);
    a.dbg_nbr_count = reinterpret_cast<uint32_t*>(base + o_cnt);
    a.knn_only = 1;
    HIP_TRY(launch_resolve(a, false, s));
    HIP_TRY(launch_knn_unpack(reinterpret_cast<const uint32_t*>(base + o_q),
                              reinterpret_cast<const int32_t*>(base + o_packed),
                              reinterpret_cast<const uint32_t*>(base + o_cnt), n_q, k, w,
                              reinterpret_cast<int32_t*>(base + o_pairs), s));
    HIP_TRY(hipMemcpyAsync(nbrs, base + o_pairs, (size_t)n_q * k * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(counts, base + o_cnt, (size_t)n_q * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return TSG_OK;
}

int tsg_random_candidates(tsg_ctx* ctx, uint32_t n_maps, const tsg_image* maps, const tsg_image* sample_masks,
                          uint64_t stage_seed, uint32_t item0, uint32_t n, uint32_t m, uint32_t* out) {
    if (!ctx || !maps || !out || n_maps == 0) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0 || m == 0) return TSG_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lk(ctx->mu);
    std::vector<std::unique_ptr<DBuf<uint8_t>>> keep;
    std::vector<MapDesc> md(n_maps);
    for (uint32_t i = 0; i < n_maps; ++i) {
        md[i] = MapDesc{};
        md[i].w = maps[i].w;
        md[i].h = maps[i].h;
        if (sample_masks && sample_masks[i].rgba) {
            std::vector<uint8_t> v((size_t)maps[i].w * maps[i].h);
            size_t valid = 0;
            for (uint32_t y = 0; y < maps[i].h; ++y)
                for (uint32_t x = 0; x < maps[i].w; ++x) {
                    uint8_t r = sample_masks[i].rgba[((size_t)y * sample_masks[i].w + x) * 4];
                    v[(size_t)y * maps[i].w + x] = r != 0;
                    valid += r != 0;
                }
            if (!valid) return fail(TSG_ERR_SAMPLING, "sample mask admits no pixel");
            keep.emplace_back(new DBuf<uint8_t>());
            HIP_TRY(keep.back()->alloc(v.size()));
            HIP_TRY(hipMemcpy(keep.back()->p, v.data(), v.size(), hipMemcpyHostToDevice));
            md[i].smask = keep.back()->p;
        }
    }
    DBuf<MapDesc> dmd;
    DBuf<uint2> dout;
    DBuf<uint32_t> ovf;
    HIP_TRY(dmd.alloc(n_maps));
    HIP_TRY(dout.alloc((size_t)n * m));
    HIP_TRY(ovf.alloc(1));
    HIP_TRY(hipMemset(ovf.p, 0, 4));
    HIP_TRY(hipMemcpy(dmd.p, md.data(), n_maps * sizeof(MapDesc), hipMemcpyHostToDevice));
    RandArgs ra{dmd.p, n_maps, stage_seed, item0, item0 + n, m, dout.p, ovf.p};
    HIP_TRY(launch_rand(ra, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<uint2> h((size_t)n * m);
    HIP_TRY(hipMemcpy(h.data(), dout.p, h.size() * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < h.size(); ++i) {
        out[i * 3] = h[i].x & 0xffffu;
        out[i * 3 + 1] = h[i].x >> 16;
        out[i * 3 + 2] = h[i].y;
    }
    return TSG_OK;
}

int tsg_score_batch(tsg_ctx* ctx, uint32_t w, uint32_t h, const uint8_t* color, const uint32_t* coord,
                    const uint32_t* ids, uint32_t n_maps, const tsg_image* maps, const tsg_image* guides,
                    const tsg_image* target_guide, const float* lut_color, const float* lut_guide, uint32_t k,
                    uint32_t cmax, uint32_t n, const uint32_t* flat, const uint32_t* nbr_count, const int32_t* nbrs,
                    const uint32_t* cand_count, const int32_t* cands, float* full_cost, uint32_t* best,
                    float* best_score, uint32_t* best_pruned, float* best_score_pruned, uint32_t team) {
    if (team > 8 || (team & (team - 1)) != 0) return fail(TSG_ERR_INVALID_ARGUMENT, "team must be 0, 1, 2, 4 or 8");
    if (!ctx || !color || !coord || !ids || !maps || !lut_color || !flat || !nbr_count || !nbrs || !cand_count ||
        !cands || !full_cost || !best || !best_score)
        return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    if (k == 0 || k > (uint32_t)kMaxK) return fail(TSG_ERR_UNSUPPORTED, "k must be in 1..64");
    bool guided = guides != nullptr;
    if (guided && (!target_guide || !lut_guide)) return fail(TSG_ERR_INVALID_ARGUMENT, "guided needs all guides");
    if (n == 0) return TSG_OK;
    for (uint32_t i = 0; i < n; ++i) {
        if (flat[i] >= w * h || nbr_count[i] > k || cand_count[i] > cmax || nbr_count[i] == 0 || cand_count[i] == 0)
            return fail(TSG_ERR_INVALID_ARGUMENT, "bad item");
        int last_kind = 0;
        uint32_t n_k0 = 0;
        for (uint32_t c = 0; c < cand_count[i]; ++c) {
            const int32_t* e = cands + ((size_t)i * cmax + c) * 6;
            if ((uint32_t)e[2] >= n_maps || e[0] < 0 || e[1] < 0 || e[0] > 65535 || e[1] > 65535)
                return fail(TSG_ERR_INVALID_ARGUMENT, "bad candidate");
            int kind = e[3] != 0;
            if (kind < last_kind) return fail(TSG_ERR_INVALID_ARGUMENT, "coherence candidates must come first");
            last_kind = kind;
            n_k0 += kind == 0;
        }
        if (n_k0 > (uint32_t)kMaxK) return fail(TSG_ERR_INVALID_ARGUMENT, "more than 64 coherence candidates");
        for (uint32_t jn = 0; jn < nbr_count[i]; ++jn) {
            int nx = nbrs[((size_t)i * k + jn) * 2], ny = nbrs[((size_t)i * k + jn) * 2 + 1];
            if (nx < 0 || ny < 0 || (uint32_t)nx >= w || (uint32_t)ny >= h)
                return fail(TSG_ERR_INVALID_ARGUMENT, "neighbour outside the image");
        }
    }
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    size_t N = (size_t)w * h;
    std::vector<uint4> st(N);
    for (size_t i = 0; i < N; ++i)
        st[i] = make_uint4(pack_px(color + i * 4), coord[i * 3] | (coord[i * 3 + 1] << 16),
                           coord[i * 3 + 2] | (ids[i * 2 + 1] << 16), ids[i * 2]);
    std::vector<std::unique_ptr<DBuf<uint32_t>>> keep;
    auto upload_img = [&](const tsg_image& im) -> uint32_t* {
        std::vector<uint32_t> v = packed(im);
        keep.emplace_back(new DBuf<uint32_t>());
        if (keep.back()->alloc(v.size()) != hipSuccess) return nullptr;
        if (hipMemcpy(keep.back()->p, v.data(), v.size() * 4, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        return keep.back()->p;
    };
    std::vector<MapDesc> md(n_maps);
    for (uint32_t i = 0; i < n_maps; ++i) {
        md[i] = MapDesc{};
        md[i].w = maps[i].w;
        md[i].h = maps[i].h;
        if (!(md[i].px = upload_img(maps[i]))) return fail(TSG_ERR_DEVICE, "upload failed");
        if (guided) {
            md[i].gw = guides[i].w;
            md[i].gh = guides[i].h;
            if (!(md[i].gpx = upload_img(guides[i]))) return fail(TSG_ERR_DEVICE, "upload failed");
        }
    }
    uint32_t* dtg = nullptr;
    if (guided && !(dtg = upload_img(*target_guide))) return fail(TSG_ERR_DEVICE, "upload failed");
    std::vector<float> lut(512, 0.0f);
    memcpy(lut.data(), lut_color, 256 * 4);
    if (guided) memcpy(lut.data() + 256, lut_guide, 256 * 4);
    std::vector<int32_t> pn((size_t)n * k, 0);
    std::vector<uint4> pc((size_t)n * cmax, make_uint4(0, 0, 0, 0));
    for (uint32_t i = 0; i < n; ++i) {
        int x = (int)(flat[i] % w), y = (int)(flat[i] / w);
        for (uint32_t jn = 0; jn < nbr_count[i]; ++jn) {
            int dx = nbrs[((size_t)i * k + jn) * 2] - x, dy = nbrs[((size_t)i * k + jn) * 2 + 1] - y;
            pn[(size_t)i * k + jn] = (int32_t)(((uint32_t)dx & 0xffffu) | ((uint32_t)dy << 16));
        }
        for (uint32_t c = 0; c < cand_count[i]; ++c) {
            const int32_t* e = cands + ((size_t)i * cmax + c) * 6;
            pc[(size_t)i * cmax + c] =
                make_uint4((uint32_t)e[0] | ((uint32_t)e[1] << 16), (uint32_t)e[2] | ((uint32_t)(e[3] != 0) << 31),
                           (uint32_t)e[4], 0u);
        }
    }
    DBuf<uint4> dst, dpc, dres;
    DBuf<MapDesc> dmd;
    DBuf<float> dlut, dfc;
    DBuf<uint32_t> dflat, dnc, dcc, dbest;
    DBuf<int32_t> dpn;
    HIP_TRY(dst.alloc(N));
    HIP_TRY(dpc.alloc(pc.size()));
    HIP_TRY(dres.alloc(n));
    HIP_TRY(dmd.alloc(n_maps));
    HIP_TRY(dlut.alloc(512));
    HIP_TRY(dfc.alloc((size_t)n * cmax));
    HIP_TRY(dflat.alloc(n));
    HIP_TRY(dnc.alloc(n));
    HIP_TRY(dcc.alloc(n));
    HIP_TRY(dbest.alloc(n));
    HIP_TRY(dpn.alloc(pn.size()));
    HIP_TRY(hipMemcpy(dst.p, st.data(), N * sizeof(uint4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dpc.p, pc.data(), pc.size() * sizeof(uint4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dmd.p, md.data(), n_maps * sizeof(MapDesc), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dlut.p, lut.data(), 512 * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dflat.p, flat, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dnc.p, nbr_count, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dcc.p, cand_count, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dpn.p, pn.data(), pn.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(dfc.p, 0, (size_t)n * cmax * 4));
    BatchArgs a{};
    a.W = w;
    a.H = h;
    a.tiles_x = (w + 7) / 8;
    a.tiles_y = (h + 7) / 8;
    a.state = dst.p;
    a.tg = dtg;
    a.maps = dmd.p;
    a.n_maps = n_maps;
    a.work = dflat.p;
    a.lut = dlut.p;
    a.item0 = 0;
    a.item1 = n;
    a.k = k;
    a.result = dres.p;
    a.ext_nbrs = dpn.p;
    a.ext_nbr_count = dnc.p;
    a.ext_cands = dpc.p;
    a.ext_cand_count = dcc.p;
    a.ext_cmax = cmax;
    a.ext_full_cost = dfc.p;
    a.dbg_best = dbest.p;
    bool lut_nan = false;
    for (float v : lut) lut_nan |= std::isnan(v);
    a.lut_nan = lut_nan;
    HIP_TRY(launch_resolve(a, guided, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<uint4> res(n);
    HIP_TRY(hipMemcpy(res.data(), dres.p, n * sizeof(uint4), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(full_cost, dfc.p, (size_t)n * cmax * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(best, dbest.p, n * 4, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) memcpy(&best_score[i], &res[i].w, 4);
    if (best_pruned || best_score_pruned) {
        // production evaluation (dedup + pruning) of the same lists
        a.ext_full_cost = nullptr;
        a.team = team;
        HIP_TRY(launch_resolve(a, guided, s));
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipMemcpy(res.data(), dres.p, n * sizeof(uint4), hipMemcpyDeviceToHost));
        std::vector<uint32_t> bp(n);
        HIP_TRY(hipMemcpy(bp.data(), dbest.p, n * 4, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) {
            if (best_pruned) best_pruned[i] = bp[i];
            if (best_score_pruned) memcpy(&best_score_pruned[i], &res[i].w, 4);
        }
    }
    return TSG_OK;
}


int tsg_job_footprint(const tsg_job* job, uint64_t* bytes) {
    if (!job || !bytes) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    *bytes = job_footprint(job);
    return TSG_OK;
}

int tsg_footprint_estimate(const tsg_params* prm, uint32_t n_examples, const uint32_t* example_wh,
                           const int* sample_kind, const uint32_t* guide_wh, uint32_t out_w, uint32_t out_h,
                           uint64_t* bytes) {
    if (!prm || !example_wh || !bytes) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    const uint64_t levels = prm->backtrack_stages == 0 ? 1 : prm->backtrack_stages;
    const uint64_t N = (uint64_t)out_w * out_h;
    const uint64_t P = prm->backtrack_stages;
    const uint64_t m = std::max<uint64_t>(1, prm->random_sample_locations);
    uint64_t b = 0, nf = 0;
    for (uint32_t i = 0; i < n_examples; ++i) {
        const uint64_t px = (uint64_t)example_wh[2 * i] * example_wh[2 * i + 1];
        const int kd = sample_kind ? sample_kind[i] : TSG_SAMPLE_ALL;
        b += levels * px * 4 + (kd == TSG_SAMPLE_IMAGE ? px : 0);
        if (guide_wh) b += levels * (uint64_t)guide_wh[2 * i] * guide_wh[2 * i + 1] * 4;
        nf += kd != TSG_SAMPLE_IGNORE;
    }
    if (guide_wh) b += levels * N * 4;
    b += levels * nf * sizeof(MapDesc);
    // state, score, resolved list, per-item scores, resolved-mask tiles (and ghost flags)
    b += N * (sizeof(uint4) + 12) + (uint64_t)((out_w + 7) / 8) * ((out_h + 7) / 8) * (prm->tiling_mode ? 16 : 8);
    // scratch under the engine: largest batch results, one random-candidate
    // chunk, per-stage statistics (default schedule: largest batch N >> 5)
    const uint64_t maxb = std::min<uint64_t>(N, std::max<uint64_t>((N >> 5) + 1, small_batch_max()));
    b += maxb * sizeof(uint4);
    b += std::min<uint64_t>(N, std::max<uint64_t>(maxb, kEngineRcBudget / (8 * m))) * m * sizeof(uint2);
    b += (P + 1) * 16 + 512 * 4 + 64;
    // engine: per-stage cost tables, stage-statistics partial sums
    b += (P + 1) * 512 * 4 + (N + kStatsChunk - 1) / kStatsChunk * 16;
    // rank map, window mask, prefetched neighbour lists
    const uint64_t win = std::min<uint64_t>(N, kPrefetchCap);
    b += N * 4 + (uint64_t)((out_w + 7) / 8) * ((out_h + 7) / 8) * 12 + win * (8 + 4ull * prm->nearest_neighbors);
    *bytes = b;
    return TSG_OK;
}

int tsg_batch_create(tsg_ctx* ctx, uint32_t max_active, tsg_batch** out) {
    if (!ctx || !out) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    if (max_active == 0) return fail(TSG_ERR_INVALID_ARGUMENT, "max_active must be positive");
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<tsg_batch> b(new tsg_batch());
    b->ctx = ctx;
    b->max_active = max_active;
    tsg_tuning_default(&b->tune);
    *out = b.release();
    return TSG_OK;
}

int tsg_batch_submit(tsg_batch* b, tsg_job* job, uint64_t* ticket) {
    if (!b || !job || !ticket) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    if (job->ctx != b->ctx) return fail(TSG_ERR_INVALID_ARGUMENT, "job belongs to another context");
    if (job->ran || job->step.active) return fail(TSG_ERR_INVALID_ARGUMENT, "a job runs once");
    job->rc_budget = kEngineRcBudget;
    job->in_engine = true;
    job->pk_threads = 2;
    for (auto* v : {&b->active, &b->pending, &b->done, &b->taken})
        for (const tsg_batch::Entry& e : *v)
            if (e.job == job) return fail(TSG_ERR_INVALID_ARGUMENT, "job already submitted");
    *ticket = b->next_ticket++;
    job->owner = b;
    b->pending.push_back({*ticket, job, 0, false});
    return TSG_OK;
}

int tsg_debug_exp_sweep(tsg_ctx* ctx, uint64_t i0, uint64_t n, float* out) {
    if (!ctx || (!out && n)) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    if (n > (1ull << 28)) return fail(TSG_ERR_INVALID_ARGUMENT, "at most 2^28 values per call");
    if (n == 0) return TSG_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lk(ctx->mu);
    DBuf<float> d;
    HIP_TRY(d.alloc(n));
    HIP_TRY(launch_exp_sweep(i0, n, d.p, ctx->stream));
    HIP_TRY(hipMemcpyAsync(out, d.p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return TSG_OK;
}

int tsg_debug_exp_sweep_host(uint64_t i0, uint64_t n, float* out) {
    if (!out && n) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    for (uint64_t r = 0; r < n; ++r) {
        const uint64_t i = i0 + r;
        const double t = (double)((i * 0x9E3779B97F4A7C15ull) >> 11) * 0x1p-53;
        const double x = (i & 3u) ? t * 8.0 : 8.0 + t * 56.0;
        out[r] = (float)exp_dd(-1.0 * x);
    }
    return TSG_OK;
}

int tsg_batch_engine_profile(tsg_batch* b, double* out11) {
    if (!b || !out11) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    for (int q = 0; q < 11; ++q) out11[q] = b->prof[q];
    return TSG_OK;
}

int tsg_batch_set_quantum(tsg_batch* b, uint64_t items) {
    if (!b) return fail(TSG_ERR_INVALID_ARGUMENT, "null batch");
    b->quantum = items;
    return TSG_OK;
}

int tsg_batch_set_alone_below(tsg_batch* b, uint32_t jobs, uint64_t ahead_bytes) {
    if (!b) return fail(TSG_ERR_INVALID_ARGUMENT, "null batch");
    b->alone_below = jobs;
    b->alone_ahead = ahead_bytes;
    return TSG_OK;
}

int tsg_batch_debug_engine(tsg_batch* b, uint32_t wgs_cap, uint32_t flags) {
    if (!b) return fail(TSG_ERR_INVALID_ARGUMENT, "null batch");
    b->wgs_cap = wgs_cap;
    b->engine_flags = flags;
    return TSG_OK;
}

int tsg_debug_fail_allocs(int64_t count) {
    g_fail_allocs.store(count < 0 ? 0 : count);
    return TSG_OK;
}

int tsg_batch_cancel(tsg_batch* b, uint64_t ticket) {
    if (!b) return fail(TSG_ERR_INVALID_ARGUMENT, "null batch");
    HIP_TRY(hipSetDevice(b->ctx->device));
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    for (size_t q = 0; q < b->pending.size(); ++q)
        if (b->pending[q].ticket == ticket) {
            tsg_batch::Entry e = b->pending[q];
            b->pending.erase(b->pending.begin() + q);
            e.status = TSG_ERR_CANCELLED;
            e.failed = true;
            b->taken.push_back(e);
            return TSG_OK;
        }
    // between launches no engine work of an active job is in flight
    for (size_t q = 0; q < b->active.size(); ++q)
        if (b->active[q].ticket == ticket) return eng_drop(b, q, TSG_ERR_CANCELLED);
    return fail(TSG_ERR_INVALID_ARGUMENT, "ticket not pending or running");
}

int tsg_batch_step(tsg_batch* b, uint32_t max_rounds, uint32_t* in_flight) {
    if (!b) return fail(TSG_ERR_INVALID_ARGUMENT, "null batch");
    HIP_TRY(hipSetDevice(b->ctx->device));
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    for (uint32_t r = 0; r < max_rounds && !(b->active.empty() && b->pending.empty()); ++r) {
        int rc = batch_round(b);
        if (rc) return rc;
    }
    int rc = batch_reap(b, false);
    if (rc) return rc;
    if (in_flight) *in_flight = (uint32_t)(b->active.size() + b->pending.size() + b->done.size());
    return TSG_OK;
}

int tsg_batch_run_all(tsg_batch* b) {
    if (!b) return fail(TSG_ERR_INVALID_ARGUMENT, "null batch");
    HIP_TRY(hipSetDevice(b->ctx->device));
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    while (!(b->active.empty() && b->pending.empty())) {
        int rc = batch_round(b);
        if (rc) return rc;
    }
    return batch_reap(b, true);
}

int tsg_batch_poll(tsg_batch* b, uint64_t* tickets, int* status, uint32_t cap, uint32_t* n) {
    if (!b || !n) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(b->ctx->device));
    {
        std::lock_guard<std::mutex> lk(b->ctx->mu);
        int rc = batch_reap(b, false);
        if (rc) return rc;
    }
    uint32_t c = 0;
    for (const tsg_batch::Entry& e : b->taken) {
        if (c >= cap) break;
        if (tickets) tickets[c] = e.ticket;
        if (status) status[c] = e.status;
        ++c;
    }
    *n = c;
    return TSG_OK;
}

int tsg_batch_take(tsg_batch* b, uint64_t ticket, tsg_job** job) {
    if (!b || !job) return fail(TSG_ERR_INVALID_ARGUMENT, "null argument");
    for (size_t q = 0; q < b->taken.size(); ++q)
        if (b->taken[q].ticket == ticket) {
            *job = b->taken[q].job;
            (*job)->owner = nullptr;
            b->taken.erase(b->taken.begin() + q);
            return TSG_OK;
        }
    return fail(TSG_ERR_INVALID_ARGUMENT, "ticket not finished or unknown");
}

int tsg_batch_stats(tsg_batch* b, uint64_t* rounds, uint64_t* launches, uint64_t* units, uint64_t* alone) {
    if (!b) return fail(TSG_ERR_INVALID_ARGUMENT, "null batch");
    if (rounds) *rounds = b->rounds;
    if (launches) *launches = b->launches;
    if (units) *units = b->units_total;
    if (alone) *alone = b->alone_runs;
    return TSG_OK;
}

void tsg_batch_destroy(tsg_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    for (auto* v : {&b->active, &b->pending, &b->done, &b->taken})
        for (auto& e : *v) {
            e.job->owner = nullptr;
            tsg_job_destroy(e.job);
        }
    delete b;
}

}  // extern "C"
