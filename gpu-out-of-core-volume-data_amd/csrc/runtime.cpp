// runtime.cpp -- host runtime of libgcvol.so: brick sources, the request
// handler, the cache manager state machine, frames, and the C ABI of
// include/gcvol.h.
//
// The cache manager follows CacheManager<T> (CacheManager.hpp:136-672):
//   update() = LRU update, then either load the bricks of the finished
//   request or compact the misses of the last frame into a new request.
// Device state (HBM):
//   mrpd     uint4 per page-directory cell (levels side by side along x)
//   request  uint32 per page-directory cell (timestamp of the last miss)
//   pool     slots x bs^3 voxels, brick-major (slot = x + cx*(y + cy*z))
//   usage    uint32 per slot (timestamp of the last frame that sampled it)
//   order    LRU order as a permutation of slot ids (most recent first)
//   meta     per slot {volumePosition, level}
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <atomic>
#include <deque>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "gc_internal.hpp"
#include "source.hpp"
#include "synth.hpp"

namespace gc {

// =================================================================== errors
static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

static uint32_t hw_threads();

// One process-wide pool of host worker threads for parallel_for.  Several
// request handlers (one per cache) can fetch bricks at once: every job sits
// in one queue that the workers serve round-robin, a few items at a time, so
// concurrent jobs share all the cores for as long as they run (a job that
// starts while another holds the cores is not left with one thread).  The
// calling thread works on its own job too, so a job always progresses.
namespace {
struct PoolJob {
    const std::function<void(uint32_t)> *fn = nullptr;
    uint32_t n = 0;
    uint32_t cap = 1;                    // most threads at once, the caller included
    std::atomic<uint32_t> next{0};       // next unclaimed item
    std::atomic<uint32_t> left{0};       // items not finished
    uint32_t workers = 0;                // pool threads inside the job (pool lock)
    bool queued = true;                  // still in the pool's queue (pool lock)
};

class WorkPool {
public:
    static WorkPool &get()
    {
        // never destroyed: its threads wait for work until the process ends,
        // so a handler still running at exit never meets a destroyed pool
        static WorkPool *pool = new WorkPool();
        return *pool;
    }

    void run(uint32_t n, uint32_t cap, const std::function<void(uint32_t)> &fn)
    {
        PoolJob j;
        j.fn = &fn;
        j.n = n;
        j.cap = cap;
        j.left = n;
        {
            std::lock_guard<std::mutex> lk(m);
            jobs.push_back(&j);
        }
        cv.notify_all();
        for (;;) {
            const uint32_t i = j.next.fetch_add(1);
            if (i >= n) break;
            fn(i);
            j.left.fetch_sub(1);
        }
        std::unique_lock<std::mutex> lk(m);
        dequeue(&j);
        done_cv.wait(lk, [&] { return j.left.load() == 0 && j.workers == 0; });
    }

private:
    static constexpr uint32_t BATCH = 4;   // items per visit to a job
    std::mutex m;
    std::condition_variable cv, done_cv;
    std::deque<PoolJob *> jobs;            // jobs with unclaimed items
    std::vector<std::thread> threads;

    WorkPool()
    {
        const uint32_t n = hw_threads();
        for (uint32_t t = 0; t < n; ++t) threads.emplace_back([this] { loop(); });
    }
    void dequeue(PoolJob *j)
    {
        if (!j->queued) return;
        jobs.erase(std::find(jobs.begin(), jobs.end(), j));
        j->queued = false;
    }
    void loop()
    {
        std::unique_lock<std::mutex> lk(m);
        for (;;) {
            PoolJob *j = nullptr;
            cv.wait(lk, [&] {
                // the next job in round-robin order that has room for a thread
                for (size_t k = 0; k < jobs.size(); ++k) {
                    PoolJob *c = jobs.front();
                    jobs.pop_front();
                    jobs.push_back(c);
                    if (c->workers + 1 < c->cap) {
                        j = c;
                        return true;
                    }
                }
                return false;
            });
            ++j->workers;
            lk.unlock();
            bool exhausted = false;
            for (uint32_t k = 0; k < BATCH; ++k) {
                const uint32_t i = j->next.fetch_add(1);
                if (i >= j->n) {
                    exhausted = true;
                    break;
                }
                (*j->fn)(i);
                j->left.fetch_sub(1);
            }
            lk.lock();
            --j->workers;
            if (exhausted) dequeue(j);
            done_cv.notify_all();
            cv.notify_all();   // a slot in this job may have freed up
        }
    }
};
} // namespace

static void parallel_for(uint32_t n, uint32_t threads, const std::function<void(uint32_t)> &fn)
{
    if (threads <= 1 || n <= 1) {
        for (uint32_t i = 0; i < n; ++i) fn(i);
        return;
    }
    WorkPool::get().run(n, std::min(threads, n), fn);
}

// parallel_for over n items in chunks of `chunk` consecutive items:
// done(c) runs once per chunk, on the thread that finishes its last item.
static void parallel_for_chunks(uint32_t n, uint32_t threads, uint32_t chunk,
                                const std::function<void(uint32_t)> &fn,
                                const std::function<void(uint32_t)> &done)
{
    const uint32_t nc = (n + chunk - 1) / chunk;
    std::unique_ptr<std::atomic<uint32_t>[]> left(new std::atomic<uint32_t>[nc]);
    for (uint32_t c = 0; c < nc; ++c) left[c] = std::min(chunk, n - c * chunk);
    parallel_for(n, threads, [&](uint32_t i) {
        fn(i);
        if (left[i / chunk].fetch_sub(1) == 1) done(i / chunk);
    });
}

static uint32_t hw_threads()
{
    const uint32_t h = std::thread::hardware_concurrency();
    return h ? h : 1;
}

// ================================================================== sources
// Host pyramid, bricks cut with the covering rule of Bricker_v2.cpp:138-222.
struct PyramidSource : Source {
    std::vector<const uint8_t *> levels;
    int get(uint32_t level, const uint32_t bbc[3], void *dst) override
    {
        if (level >= vc.nb_levels) return GC_BRICK_ERROR;
        const uint32_t *d = vc.initial_sizes[level];
        const uint8_t *src = levels[level];
        const uint32_t inner = bs - 2 * cov;
        const uint64_t BX = (uint64_t)bs * BB[0], BY = (uint64_t)bs * BB[1];
        std::memset(dst, 0, bb_bytes);
        uint8_t *out = static_cast<uint8_t *>(dst);
        for (uint32_t sz = 0; sz < BB[2]; ++sz)
            for (uint32_t sy = 0; sy < BB[1]; ++sy)
                for (uint32_t sx = 0; sx < BB[0]; ++sx) {
                    const int64_t ox = (int64_t)(bbc[0] * BB[0] + sx) * inner - cov;
                    const int64_t oy = (int64_t)(bbc[1] * BB[1] + sy) * inner - cov;
                    const int64_t oz = (int64_t)(bbc[2] * BB[2] + sz) * inner - cov;
                    const int64_t x0 = std::max<int64_t>(ox, 0), x1 = std::min<int64_t>(ox + bs, d[0]);
                    if (x1 <= x0) continue;
                    for (uint32_t k = 0; k < bs; ++k) {
                        const int64_t gz = oz + k;
                        if (gz < 0 || gz >= d[2]) continue;
                        for (uint32_t j = 0; j < bs; ++j) {
                            const int64_t gy = oy + j;
                            if (gy < 0 || gy >= d[1]) continue;
                            const uint64_t so = ((uint64_t)x0 + (uint64_t)d[0] * ((uint64_t)gy + (uint64_t)d[1] * gz)) * bytes;
                            const uint64_t dof = (sx * (uint64_t)bs + (uint64_t)(x0 - ox)) +
                                                 BX * ((sy * (uint64_t)bs + j) + BY * (sz * (uint64_t)bs + k));
                            std::memcpy(out + dof * bytes, src + so, (size_t)(x1 - x0) * bytes);
                        }
                    }
                }
        return GC_BRICK_SUCCESS;
    }
};

struct SynthSource : Source {
    std::unique_ptr<Synth> synth;
    uint32_t nthreads = 1;
    uint32_t threads() const override { return nthreads; }
    int get(uint32_t level, const uint32_t bbc[3], void *dst) override
    {
        if (level >= vc.nb_levels) return GC_BRICK_ERROR;
        const uint32_t inner = bs - 2 * cov;
        const uint32_t n[3] = {bs, bs, bs};
        if (BB[0] == 1 && BB[1] == 1 && BB[2] == 1) {
            std::memset(dst, 0, bb_bytes);
            const int64_t o[3] = {(int64_t)bbc[0] * inner - cov, (int64_t)bbc[1] * inner - cov,
                                  (int64_t)bbc[2] * inner - cov};
            synth->fill_box(level, o, n, vc.initial_sizes[level], dst);
            return GC_BRICK_SUCCESS;
        }
        std::vector<uint8_t> tmp((size_t)bs * bs * bs * bytes);
        const uint64_t BX = (uint64_t)bs * BB[0], BY = (uint64_t)bs * BB[1];
        for (uint32_t sz = 0; sz < BB[2]; ++sz)
            for (uint32_t sy = 0; sy < BB[1]; ++sy)
                for (uint32_t sx = 0; sx < BB[0]; ++sx) {
                    std::fill(tmp.begin(), tmp.end(), 0);
                    const int64_t o[3] = {(int64_t)(bbc[0] * BB[0] + sx) * inner - cov,
                                          (int64_t)(bbc[1] * BB[1] + sy) * inner - cov,
                                          (int64_t)(bbc[2] * BB[2] + sz) * inner - cov};
                    synth->fill_box(level, o, n, vc.initial_sizes[level], tmp.data());
                    for (uint32_t k = 0; k < bs; ++k)
                        for (uint32_t j = 0; j < bs; ++j) {
                            const uint64_t dof = (uint64_t)sx * bs + BX * ((uint64_t)sy * bs + j + BY * ((uint64_t)sz * bs + k));
                            std::memcpy(static_cast<uint8_t *>(dst) + dof * bytes,
                                        tmp.data() + ((uint64_t)bs * (j + (uint64_t)bs * k)) * bytes, (size_t)bs * bytes);
                        }
                }
        return GC_BRICK_SUCCESS;
    }
};

// Synthetic volume written on the device (k_synth_bricks): the host's axis
// tables of every level go to the device once; requested bricks are
// synthesised straight into device memory, bit-identical to SynthSource's.
// get() (host bricks) stays available.
struct DeviceSynthSource : SynthSource {
    int device = 0;
    float *d_tab = nullptr, *d_amp = nullptr;
    uint32_t *d_keys = nullptr, *d_below = nullptr;
    uint32_t cap_n = 0;
    SynthLaunch base{};
    std::mutex m;
    bool on_device() const override { return true; }
    ~DeviceSynthSource() override
    {
        hipFree(d_tab); hipFree(d_amp); hipFree(d_keys); hipFree(d_below);
    }
    int setup()
    {
        const Synth &S = *synth;
        const uint32_t nb = (uint32_t)S.blobs.size();
        if (BB[0] != 1 || BB[1] != 1 || BB[2] != 1)
            return fail(GC_ERR_CONFIG, "bricks synthesised on the device need BigBrickSize 1");
        if ((uint64_t)3 * nb * bs * sizeof(float) > 65536)
            return fail(GC_ERR_CONFIG, "3 x blobs x brick size factors exceed 64 KiB of LDS");
        if (vc.nb_levels == 0 || vc.nb_levels > GC_MAX_LEVELS) return fail(GC_ERR_CONFIG, "bad level count");
        // factors of every level over its extent plus a margin on each side
        // (covering voxels of border bricks lie outside the extent and are 0)
        const int64_t pad = Synth::TABLE_PAD;
        std::vector<float> tab;
        for (uint32_t l = 0; l < vc.nb_levels; ++l)
            for (int a = 0; a < 3; ++a) {
                SynthAxis &ax = base.axis[l][a];
                ax.off = tab.size();
                ax.lo = -pad;
                ax.len = vc.initial_sizes[l][a] + 2 * (uint32_t)pad;
                tab.resize(tab.size() + (size_t)nb * ax.len);
                for (uint32_t k = 0; k < nb; ++k)
                    S.axis_factors(k, a, l, ax.lo, ax.len, 1.f, &tab[ax.off + (size_t)k * ax.len]);
                base.ext[l][a] = vc.initial_sizes[l][a];
            }
        std::vector<float> amp(nb);
        for (uint32_t k = 0; k < nb; ++k) amp[k] = (float)S.blobs[k].amp;
        GC_HIP(hipSetDevice(device));
        GC_HIP(hipMalloc((void **)&d_tab, std::max<size_t>(tab.size(), 1) * sizeof(float)));
        GC_HIP(hipMalloc((void **)&d_amp, std::max<size_t>(nb, 1) * sizeof(float)));
        GC_HIP(hipMemcpy(d_tab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
        GC_HIP(hipMemcpy(d_amp, amp.data(), nb * sizeof(float), hipMemcpyHostToDevice));
        base.tab = d_tab;
        base.amp = d_amp;
        base.nb = nb;
        base.bs = bs;
        base.cov = cov;
        base.bytes = bytes;
        base.seed = (uint32_t)S.prm.seed;
        base.noise = S.prm.noise;
        base.fmax = (float)S.maxval;
        base.max_level = ~0u;
        return GC_OK;
    }
    // Bricks of the keys that the synthesis kernel skipped (coarser levels
    // of a source that builds them otherwise), on stream s.
    virtual int write_other(uint32_t n, uint8_t *d_dst, float thr, hipStream_t s)
    {
        (void)n; (void)d_dst; (void)thr; (void)s;
        return GC_OK;
    }
    int get_device(const uint32_t *keys, uint32_t n, uint8_t *d_dst, float thr, uint8_t *below,
                   hipStream_t s) override
    {
        if (n == 0) return GC_OK;
        int cur = -1;
        GC_HIP(hipGetDevice(&cur));
        if (cur != device) return fail(GC_ERR_ARG, "brick source lives on another device");
        for (uint32_t i = 0; i < n; ++i)
            if (keys[4 * i + 3] >= vc.nb_levels) return fail(GC_ERR_SOURCE, "brick level out of range");
        std::lock_guard<std::mutex> lk(m);
        if (n > cap_n) {
            hipFree(d_keys);
            hipFree(d_below);
            d_keys = d_below = nullptr;
            cap_n = 0;
            GC_HIP(hipMalloc((void **)&d_keys, (size_t)n * 16));
            GC_HIP(hipMalloc((void **)&d_below, (size_t)n * 4));
            cap_n = n;
        }
        std::vector<uint32_t> flags(n);
        SynthLaunch p = base;
        p.keys = d_keys;
        p.dst = d_dst;
        p.below = d_below;
        p.n = n;
        p.thr = thr;
        GC_HIP(hipMemcpyAsync(d_keys, keys, (size_t)n * 16, hipMemcpyHostToDevice, s));
        GC_HIP(launch_synth_bricks(p, s));
        GC_TRY(write_other(n, d_dst, thr, s));
        GC_HIP(hipMemcpyAsync(flags.data(), d_below, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        GC_HIP(hipStreamSynchronize(s));
        for (uint32_t i = 0; i < n; ++i) below[i] = (uint8_t)flags[i];
        return GC_OK;
    }
};

// The synthetic volume's level 0 with a resident mip pyramid on the device
// (the preprocessing of Mipmapper.cu:401-597 for a volume that does not fit
// anywhere at level 0): a cold build synthesises level 0 in z-slabs of
// `slab` planes straight into device memory (k_synth_box_r4, the voxels of the
// brick synthesiser) and reduces each slab into level 1 (k_mip2, the
// Mipmapper rule; slabs start on even planes, so each 2x2x2 group lies in one
// slab); levels 2.. are reduced from the resident level 1.  Requests then
// synthesise level-0 bricks and cut coarser bricks from the resident levels
// with the covering rule, both on the device.  The coarse levels are the
// CPU preprocessor's bit for bit; the level-0 volume is never stored.
struct DevicePyramidSource : DeviceSynthSource {
    std::vector<uint8_t *> d_levels;   // [l] for l >= 1
    CutKeys cut{};
    double build_ms = 0.0;
    ~DevicePyramidSource() override
    {
        for (uint8_t *p : d_levels) hipFree(p);
    }
    int build(uint32_t slab)
    {
        const uint32_t nl = vc.nb_levels;
        if (slab == 0 || slab % 2) return fail(GC_ERR_ARG, "slab planes must be even and positive");
        for (uint32_t l = 1; l < nl; ++l)
            for (int a = 0; a < 3; ++a)
                if (vc.initial_sizes[l][a] != vc.initial_sizes[l - 1][a] / 2)
                    return fail(GC_ERR_CONFIG, "level sizes do not halve (floor) from level to level");
        const uint32_t *d0 = vc.initial_sizes[0];
        if (d0[1] > 65535u) return fail(GC_ERR_CONFIG, "level 0 taller than 65535 voxels");
        GC_HIP(hipSetDevice(device));
        const auto t0 = std::chrono::steady_clock::now();
        d_levels.assign(nl, nullptr);
        for (uint32_t l = 1; l < nl; ++l) {
            const uint64_t n = (uint64_t)vc.initial_sizes[l][0] * vc.initial_sizes[l][1] * vc.initial_sizes[l][2] * bytes;
            GC_HIP(hipMalloc((void **)&d_levels[l], std::max<uint64_t>(n, 1)));
        }
        const uint64_t plane = (uint64_t)d0[0] * d0[1] * bytes;
        uint8_t *d_slab = nullptr;
        GC_HIP(hipMalloc((void **)&d_slab, plane * std::min<uint32_t>(slab, d0[2])));
        hipStream_t s = nullptr;
        hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        const uint64_t plane1 = nl > 1 ? (uint64_t)vc.initial_sizes[1][0] * vc.initial_sizes[1][1] * bytes : 0;
        for (uint32_t z0 = 0; e == hipSuccess && z0 < d0[2]; z0 += slab) {
            const uint32_t T = std::min<uint32_t>(slab, d0[2] - z0);
            SynthBox b{};
            b.tab = base.tab;
            b.amp = base.amp;
            b.dst = d_slab;
            b.o[0] = 0; b.o[1] = 0; b.o[2] = z0;
            b.n[0] = d0[0]; b.n[1] = d0[1]; b.n[2] = T;
            b.level = 0;
            b.nb = base.nb;
            b.bytes = bytes;
            b.seed = base.seed;
            b.noise = base.noise;
            b.fmax = base.fmax;
            for (int a = 0; a < 3; ++a) {
                b.axis[a] = base.axis[0][a];
                b.ext[a] = base.ext[0][a];
            }
            e = launch_synth_box(b, s);
            if (e == hipSuccess && nl > 1) {
                const uint32_t dims[3] = {d0[0], d0[1], T};
                e = launch_mip(d_slab, dims, bytes, d_levels[1] + (uint64_t)(z0 / 2) * plane1, s);
            }
        }
        for (uint32_t l = 2; e == hipSuccess && l < nl; ++l)
            e = launch_mip(d_levels[l - 1], vc.initial_sizes[l - 1], bytes, d_levels[l], s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (s) hipStreamDestroy(s);
        hipFree(d_slab);
        GC_HIP(e);
        build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (uint32_t l = 1; l < nl; ++l) {
            cut.levels[l] = d_levels[l];
            for (int a = 0; a < 3; ++a) cut.dims[l][a] = vc.initial_sizes[l][a];
        }
        cut.bs = bs;
        cut.cov = cov;
        cut.bytes = bytes;
        cut.first_level = 1;
        base.max_level = 0;
        return GC_OK;
    }
    int write_other(uint32_t n, uint8_t *d_dst, float thr, hipStream_t s) override
    {
        CutKeys c = cut;
        c.keys = d_keys;
        c.out = d_dst;
        c.below = d_below;
        c.n = n;
        c.thr = thr;
        GC_HIP(launch_cut_keys(c, s));
        return GC_OK;
    }
    // host bricks (tests / tools): level 0 from the host synthesiser (the
    // same voxels), coarser levels cut on the device and copied back
    int get(uint32_t level, const uint32_t bbc[3], void *dst) override
    {
        if (level >= vc.nb_levels) return GC_BRICK_ERROR;
        if (level == 0) return SynthSource::get(level, bbc, dst);
        if (hipSetDevice(device) != hipSuccess) return GC_BRICK_ERROR;
        const uint32_t key[4] = {bbc[0], bbc[1], bbc[2], level};
        uint8_t *d = nullptr;
        if (hipMalloc((void **)&d, bb_bytes) != hipSuccess) return GC_BRICK_ERROR;
        hipStream_t s = nullptr;
        uint8_t below = 0;
        int rc = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess ? get_device(key, 1, d, 0.f, &below, s)
                                                                                 : GC_ERR_HIP;
        if (rc == GC_OK && hipMemcpyAsync(dst, d, bb_bytes, hipMemcpyDeviceToHost, s) != hipSuccess) rc = GC_ERR_HIP;
        if (rc == GC_OK && hipStreamSynchronize(s) != hipSuccess) rc = GC_ERR_HIP;
        if (s) hipStreamDestroy(s);
        hipFree(d);
        return rc == GC_OK ? GC_BRICK_SUCCESS : GC_BRICK_ERROR;
    }
};

// Coarse bricks built on the GPU on request (mip on miss): level-0 bricks
// come from a base source; a level-l brick is reduced on the device from the
// level-(l-1) bricks whose inner regions hold its 2x2x2 groups, which are
// obtained the same way, recursively, and remembered in a bounded host cache
// (least recently used out).  The result is the brick the CPU preprocessor
// would cut from its mip level: floor(mean) with the Mipmapper's accumulator
// types, covering voxels included, zero outside the level.
struct MipSource : Source {
    Source *base = nullptr;     // borrowed
    int device = 0;
    uint64_t memo_cap = 0, memo_bytes = 0, built = 0;
    std::mutex memo_m, gpu_m;
    using Key = std::array<uint32_t, 4>;
    std::list<Key> lru;
    std::map<Key, std::pair<std::vector<uint8_t>, std::list<Key>::iterator>> memo;
    hipStream_t stream = nullptr;
    uint8_t *d_children = nullptr, *d_out = nullptr;
    uint64_t d_cap = 0;
    bool thread_safe() const override { return base->thread_safe(); }
    uint32_t threads() const override { return base->threads(); }
    ~MipSource() override
    {
        hipFree(d_children);
        hipFree(d_out);
        if (stream) hipStreamDestroy(stream);
    }
    bool lookup(const Key &k, void *dst)
    {
        std::lock_guard<std::mutex> lk(memo_m);
        auto it = memo.find(k);
        if (it == memo.end()) return false;
        std::memcpy(dst, it->second.first.data(), it->second.first.size());
        lru.splice(lru.begin(), lru, it->second.second);
        return true;
    }
    void remember(const Key &k, const void *src, size_t n)
    {
        if (memo_cap == 0 || n > memo_cap) return;
        std::lock_guard<std::mutex> lk(memo_m);
        if (memo.count(k)) return;
        lru.push_front(k);
        auto &e = memo[k];
        e.first.assign(static_cast<const uint8_t *>(src), static_cast<const uint8_t *>(src) + n);
        e.second = lru.begin();
        memo_bytes += n;
        while (memo_bytes > memo_cap && !lru.empty()) {
            auto old = memo.find(lru.back());
            memo_bytes -= old->second.first.size();
            memo.erase(old);
            lru.pop_back();
        }
    }
    int brick(uint32_t level, const uint32_t b[3], void *dst)
    {
        if (level == 0) {
            const int st = base->get(0, b, dst);
            if (st == GC_BRICK_EMPTY) std::memset(dst, 0, bb_bytes);
            return st == GC_BRICK_ERROR ? st : GC_BRICK_SUCCESS;
        }
        const Key key{level, b[0], b[1], b[2]};
        if (lookup(key, dst)) return GC_BRICK_SUCCESS;
        const uint32_t inner = bs - 2 * cov;
        const uint32_t *L = vc.initial_sizes[level];
        MipBrickArgs a{};
        a.bs = bs;
        a.cov = cov;
        a.inner = inner;
        uint64_t nchild = 1;
        for (int k = 0; k < 3; ++k) {
            a.g0[k] = (int64_t)b[k] * inner - cov;
            a.dims[k] = L[k];
            const int64_t lo = std::max<int64_t>(a.g0[k], 0), hi = std::min<int64_t>(a.g0[k] + bs - 1, (int64_t)L[k] - 1);
            if (lo > hi) {   // the brick lies outside the level
                std::memset(dst, 0, bb_bytes);
                return GC_BRICK_SUCCESS;
            }
            a.c0[k] = (uint32_t)(2 * lo / inner);
            a.nc[k] = (uint32_t)((2 * hi + 1) / inner) - a.c0[k] + 1;
            nchild *= a.nc[k];
        }
        std::vector<uint8_t> children(nchild * bb_bytes);
        for (uint32_t z = 0; z < a.nc[2]; ++z)
            for (uint32_t y = 0; y < a.nc[1]; ++y)
                for (uint32_t x = 0; x < a.nc[0]; ++x) {
                    const uint32_t c[3] = {a.c0[0] + x, a.c0[1] + y, a.c0[2] + z};
                    const uint64_t i = x + (uint64_t)a.nc[0] * (y + (uint64_t)a.nc[1] * z);
                    const int st = brick(level - 1, c, children.data() + i * bb_bytes);
                    if (st == GC_BRICK_ERROR) return st;
                }
        {
            std::lock_guard<std::mutex> lk(gpu_m);
            if (hipSetDevice(device) != hipSuccess) return GC_BRICK_ERROR;
            if (!stream && hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return GC_BRICK_ERROR;
            if (children.size() > d_cap) {
                hipFree(d_children);
                d_children = nullptr;
                if (hipMalloc((void **)&d_children, children.size()) != hipSuccess) return GC_BRICK_ERROR;
                d_cap = children.size();
            }
            if (!d_out && hipMalloc((void **)&d_out, bb_bytes) != hipSuccess) return GC_BRICK_ERROR;
            if (hipMemcpyAsync(d_children, children.data(), children.size(), hipMemcpyHostToDevice, stream) != hipSuccess ||
                launch_mip_brick(a, bytes, d_children, d_out, stream) != hipSuccess ||
                hipMemcpyAsync(dst, d_out, bb_bytes, hipMemcpyDeviceToHost, stream) != hipSuccess ||
                hipStreamSynchronize(stream) != hipSuccess)
                return GC_BRICK_ERROR;
            ++built;
        }
        remember(key, dst, bb_bytes);
        return GC_BRICK_SUCCESS;
    }
    int get(uint32_t level, const uint32_t bbc[3], void *dst) override
    {
        if (level >= vc.nb_levels) return GC_BRICK_ERROR;
        return brick(level, bbc, dst);
    }
};

struct CallbackSource : Source {
    gc_brick_fn fn = nullptr;
    void *user = nullptr;
    bool thread_safe() const override { return false; }
    int get(uint32_t level, const uint32_t bbc[3], void *dst) override
    {
        return fn(user, level, bbc[0], bbc[1], bbc[2], dst);
    }
};

} // namespace gc

namespace gc {

// ==================================================================== cache
static bool validate(const gc_volume_cfg &vc, const gc_cache_cfg &cc, std::string &why)
{
    if (vc.nb_levels == 0 || vc.nb_levels > GC_MAX_LEVELS) { why = "nb_levels out of range"; return false; }
    if (vc.brick_size[0] != vc.brick_size[1] || vc.brick_size[0] != vc.brick_size[2]) { why = "BrickSize must be cubic"; return false; }
    if (vc.covering[0] != vc.covering[1] || vc.covering[0] != vc.covering[2]) { why = "VoxelCovering must be uniform"; return false; }
    if (vc.brick_size[0] <= 2 * vc.covering[0]) { why = "BrickSize must exceed 2*VoxelCovering"; return false; }
    if (vc.encoded_bytes != 1 && vc.encoded_bytes != 2 && vc.encoded_bytes != 4) { why = "NumberEncodedBytes must be 1, 2 or 4"; return false; }
    if (vc.channels != 1) { why = "NumberChannels must be 1"; return false; }
    for (int a = 0; a < 3; ++a) {
        if (vc.big_brick_size[a] == 0) { why = "BigBrickSize must be >= 1"; return false; }
        if (cc.cache_size[a] == 0) { why = "CacheSize must be >= 1"; return false; }
        if (cc.block_size[a] != vc.brick_size[a]) { why = "BlockSize[0] must equal BrickSize"; return false; }
    }
    for (uint32_t l = 0; l < vc.nb_levels; ++l)
        for (int a = 0; a < 3; ++a)
            if (vc.initial_sizes[l][a] == 0 || vc.real_sizes[l][a] == 0) { why = "zero level size"; return false; }
    if (cc.nb_tables > GC_MAX_TABLES) { why = "at most GC_MAX_TABLES page-table caches"; return false; }
    for (uint32_t k = 0; k < cc.nb_tables; ++k)
        for (int a = 0; a < 3; ++a)
            if (cc.table_cache_size[k][a] == 0 || cc.table_block_size[k][a] == 0) {
                why = "page-table cache sizes and block sizes must be >= 1";
                return false;
            }
    return true;
}

static void compute_geometry(const gc_volume_cfg &vc, const gc_cache_cfg &cc, gc_geometry &g)
{
    // CacheManager.hpp:157-285: realNumberOfEntries bottom-up and
    // virtualizedNumberOfEntries top-down over the hierarchy (index 0 the
    // voxels, 1 the bricks, ..., T + 1 the MRPD), with the reference's
    // float / double rounding.
    std::memset(&g, 0, sizeof(g));
    g.nb_levels = vc.nb_levels;
    g.brick_size = vc.brick_size[0];
    g.covering = vc.covering[0];
    g.inner = g.brick_size - 2 * g.covering;
    g.encoded_bytes = vc.encoded_bytes;
    const uint32_t T = cc.nb_tables, H = T + 2;
    const double inner = (double)g.brick_size - 2.0 * (double)g.covering;
    uint32_t real[GC_MAX_TABLES + 2][GC_MAX_LEVELS][3] = {}, virt[GC_MAX_TABLES + 2][GC_MAX_LEVELS][3] = {};
    for (uint32_t l = 0; l < vc.nb_levels; ++l)
        for (int a = 0; a < 3; ++a) {
            real[0][l][a] = vc.real_sizes[l][a];
            real[1][l][a] = (uint32_t)std::ceil((double)(float)vc.real_sizes[l][a] / inner);
            for (uint32_t h = 2; h < H; ++h)
                real[h][l][a] = (uint32_t)std::ceil((float)real[h - 1][l][a] / (float)cc.table_block_size[h - 2][a]);
            virt[H - 1][l][a] = real[H - 1][l][a];
            for (uint32_t h = H - 2; h >= 1; --h)
                virt[h][l][a] = (uint32_t)std::ceil((float)virt[h + 1][l][a] * (float)cc.table_block_size[h - 1][a]);
            virt[0][l][a] = (uint32_t)std::ceil((double)(float)virt[1][l][a] * inner);
        }
    for (uint32_t l = 0; l < vc.nb_levels; ++l)
        for (int a = 0; a < 3; ++a) {
            g.level_dim[l][a] = real[1][l][a];
            g.virt_vox[l][a] = virt[0][l][a];
            g.rdv[l][a] = (float)vc.real_sizes[l][a] / (float)g.virt_vox[l][a];
            g.ior[l][a] = (float)vc.initial_sizes[l][a] / (float)vc.real_sizes[l][a];
            g.top_dim[l][a] = real[H - 1][l][a];
        }
    g.level_coord[0] = 0;
    g.top_coord[0] = 0;
    for (uint32_t l = 1; l < vc.nb_levels; ++l) {
        g.level_coord[l] = g.level_coord[l - 1] + g.level_dim[l - 1][0];
        g.top_coord[l] = g.top_coord[l - 1] + g.top_dim[l - 1][0];
    }
    for (uint32_t l = 0; l < vc.nb_levels; ++l) {
        g.mrpd_dim[0] += g.level_dim[l][0];
        g.mrpd_dim[1] = std::max(g.mrpd_dim[1], g.level_dim[l][1]);
        g.mrpd_dim[2] = std::max(g.mrpd_dim[2], g.level_dim[l][2]);
        g.top_cells_dim[0] += g.top_dim[l][0];
        g.top_cells_dim[1] = std::max(g.top_cells_dim[1], g.top_dim[l][1]);
        g.top_cells_dim[2] = std::max(g.top_cells_dim[2], g.top_dim[l][2]);
    }
    for (int a = 0; a < 3; ++a) {
        g.one_over[a] = 1.f / (float)(vc.brick_size[a] - 2 * vc.covering[a]);
        g.cache_size[a] = cc.cache_size[a];
    }
    g.nslots = cc.cache_size[0] * cc.cache_size[1] * cc.cache_size[2];
    g.mrpd_cells = (uint64_t)g.mrpd_dim[0] * g.mrpd_dim[1] * g.mrpd_dim[2];
    g.top_cells = (uint64_t)g.top_cells_dim[0] * g.top_cells_dim[1] * g.top_cells_dim[2];
    // _tableCaches[i] is hierarchy index H - 2 - i (CacheManager.hpp:223-233)
    g.nb_tables = T;
    for (uint32_t i = 0; i < T; ++i) {
        const uint32_t h = H - 2 - i;
        gc_table_geometry &t = g.table[i];
        for (int a = 0; a < 3; ++a) {
            t.block[a] = cc.table_block_size[h - 1][a];
            t.blocks[a] = cc.table_cache_size[h - 1][a];
            t.surf[a] = t.block[a] * t.blocks[a];
            for (uint32_t l = 0; l < vc.nb_levels; ++l) {
                t.virt[l][a] = virt[h][l][a];
                t.real[l][a] = real[h][l][a];
            }
        }
        t.nblocks = t.blocks[0] * t.blocks[1] * t.blocks[2];
    }
}

struct Cache;

// Request handler (RequestHandler.hpp:30-374): a worker thread that turns a
// miss list into bricks in pinned memory and uploads them on a side stream.
struct Handler {
    Cache *c = nullptr;
    std::thread worker;
    std::mutex m;
    std::condition_variable cv;
    std::condition_variable idle_cv;   // signalled when a request has been handled
    bool run = false;
    std::atomic<bool> new_request{false};
    std::vector<uint32_t> incoming;   // uint4 list as compacted
    int status = GC_OK;
    std::string error;
    // results of the last handled request
    uint32_t n = 0, nne = 0, nuniq = 0;
    bool upload_pending = false;

    void start();
    void stop();
    void notify(std::vector<uint32_t> &&list);
    int handle(const std::vector<uint32_t> &list);
    void loop();
};

struct Cache {
    gc_volume_cfg vc{};
    gc_cache_cfg cc{};
    gc_geometry g{};
    Source *src = nullptr;
    int device = 0;
    hipStream_t stream = nullptr, upload = nullptr;
    hipEvent_t uploaded = nullptr, consumed = nullptr;
    uint32_t ts = 1;
    uint32_t cap = 50;                  // request buffers' capacity (max_requests at create)
    uint32_t ray_list_max = RAY_G4_MAX;   // resumed-ray launches up to this many rays use the group-per-ray march
    uint32_t ray_group = 0;               // its lanes per ray (0: from the list length)
    // lane-march launches of the frame loop hand each ray that has consumed
    // handoff_after samples over to a group march of handoff_group lanes per
    // ray at its next mapped brick, once handoff_exited_pct % of the launch's
    // waves have run out of rays (handoff_after = 0: never).  Measured on the
    // 2048^3 u16 fly-through: 32 / 8 / 75 hands about 3 % of the rays over
    // and cuts the lane launches' time by a third (60-85 % and 8-64 samples
    // within 2 %).
    uint32_t handoff_after = 32, handoff_group = 8, handoff_exited_pct = 75;
    uint32_t group_for(uint32_t rays, bool tables) const
    {
        const uint32_t g = ray_group ? ray_group : dvr_ray_group(rays);
        return tables && g < 16u ? 16u : g;
    }
    uint32_t limit = 50;                // bricks per update, <= cap
    // device state
    uint4 *mrpd = nullptr;
    uint32_t *cells = nullptr;          // cell table: slot << 2 | w per MRPD cell (the ray caster's copy)
    uint32_t *request = nullptr, *usage = nullptr, *order = nullptr, *order_tmp = nullptr;
    SlotMeta *meta = nullptr;
    uint8_t *pool = nullptr;
    gc_geometry *d_geom = nullptr;
    LruScratch lru_s{};
    MissScratch miss_s{};
    uint32_t *d_misses = nullptr;       // cap x uint4
    uint32_t *d_count = nullptr;        // scratch counter
    // page-table caches (g.nb_tables of them)
    TableState tabs{};
    uint32_t *tab_order_tmp[GC_MAX_TABLES] = {};
    uint32_t *d_unref = nullptr;        // bricks the last load left unreferenced
    RefScratch refs{};                  // parallel referencing (one table)
    uint32_t *h_ref = nullptr;          // pinned: cap x (cell, loc, rep)
    // host staging
    uint32_t *h_misses = nullptr;       // pinned: [total, pad x3, cap x uint4]
    uint8_t *h_staging = nullptr;       // pinned: cap big bricks
    uint8_t *d_staging = nullptr;
    uint32_t *h_meta = nullptr;         // pinned: asked (cap x4) | bb_index (cap) | bb_coord (cap x4)
    uint32_t *d_meta = nullptr;
    Handler handler;
    std::mutex done_m;
    bool request_done = false;
    gc_cache_counters counters{};
    std::atomic<uint64_t> fetch_ns{0};   // the handler thread's fetch time
    // frame driver
    gc_render_stats *d_frame_stats = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    gc_launch_record *log = nullptr;     // per-launch log of gc_render_frame, or null
    uint32_t log_cap = 0, *log_count = nullptr;

    ~Cache() { release(); }
    void release();
    int create();
    int lru_update();
    int prepare_bricks(bool &done);
    int load_new_bricks();
    int update(int *done);
    int insert_serial_needed(const uint32_t *asked, uint32_t n, int &serial) const;
    // insert_serial_needed's scratch: one bit per page-directory cell, all
    // clear between calls, and the cells it set
    mutable std::vector<uint64_t> cell_bits;
    mutable std::vector<uint64_t> cells_set;
    int reference_plan(const uint32_t *asked, uint32_t n, bool &parallel);
};

void Cache::release()
{
    handler.stop();
    if (stream) hipStreamSynchronize(stream);
    if (upload) hipStreamSynchronize(upload);
    hipFree(mrpd); hipFree(cells); hipFree(request); hipFree(usage); hipFree(order); hipFree(order_tmp);
    hipFree(meta); hipFree(pool); hipFree(d_geom);
    hipFree(lru_s.masks); hipFree(lru_s.counts); hipFree(lru_s.total);
    hipFree(miss_s.masks); hipFree(miss_s.counts); hipFree(miss_s.total);
    hipFree(d_misses); hipFree(d_count); hipFree(d_staging); hipFree(d_meta);
    for (int i = 0; i < GC_MAX_TABLES; ++i) {
        hipFree(tabs.words[i]); hipFree(tabs.usage[i]); hipFree(tabs.order[i]); hipFree(tabs.meta[i]);
        hipFree(tab_order_tmp[i]);
        tabs.words[i] = tabs.usage[i] = tabs.order[i] = tab_order_tmp[i] = nullptr;
        tabs.meta[i] = nullptr;
    }
    hipFree(d_unref);
    d_unref = nullptr;
    hipFree(refs.cell); hipFree(refs.w0); hipFree(refs.inv); hipFree(refs.flag);
    refs = RefScratch{};
    if (h_ref) hipHostFree(h_ref);
    h_ref = nullptr;
    if (h_misses) hipHostFree(h_misses);
    if (h_staging) hipHostFree(h_staging);
    if (h_meta) hipHostFree(h_meta);
    if (uploaded) hipEventDestroy(uploaded);
    if (consumed) hipEventDestroy(consumed);
    if (ev0) hipEventDestroy(ev0);
    if (ev1) hipEventDestroy(ev1);
    hipFree(d_frame_stats);
    d_frame_stats = nullptr;
    ev0 = ev1 = nullptr;
    if (stream) hipStreamDestroy(stream);
    if (upload) hipStreamDestroy(upload);
    mrpd = nullptr; cells = nullptr; request = usage = order = order_tmp = nullptr; meta = nullptr; pool = nullptr;
    d_geom = nullptr; h_misses = nullptr; h_staging = nullptr; h_meta = nullptr;
    uploaded = consumed = nullptr; stream = upload = nullptr;
}

template <typename T>
static int dalloc(T **p, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) return fail(GC_ERR_OOM, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    return GC_OK;
}

template <typename T>
static int halloc(T **p, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    hipError_t e = hipHostMalloc((void **)p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return fail(GC_ERR_OOM, std::string("hipHostMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    return GC_OK;
}

int Cache::create()
{
    GC_HIP(hipSetDevice(device));
    compute_geometry(vc, cc, g);
    if (g.mrpd_cells >= (1ull << 32)) return fail(GC_ERR_CONFIG, "page directory exceeds 2^32 cells");
    if (g.nslots >= (1u << 24)) return fail(GC_ERR_CONFIG, "pool exceeds 2^24 brick slots");
    for (uint32_t i = 0; i < g.nb_tables; ++i)
        if ((uint64_t)g.table[i].nblocks * table_b3(g.table[i]) >= (1ull << 32) || g.table[i].nblocks >= (1u << 28))
            return fail(GC_ERR_CONFIG, "page-table cache exceeds 2^32 entries");
    cap = cc.max_requests ? cc.max_requests : 50;
    cap = std::min(cap, g.nslots);
    limit = cap;
    GC_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    GC_HIP(hipStreamCreateWithFlags(&upload, hipStreamNonBlocking));
    GC_HIP(hipEventCreateWithFlags(&uploaded, hipEventDisableTiming));
    GC_HIP(hipEventCreateWithFlags(&consumed, hipEventDisableTiming));
    const uint64_t bvox = (uint64_t)g.brick_size * g.brick_size * g.brick_size;
    GC_TRY(dalloc(&mrpd, g.top_cells * sizeof(uint4)));
    GC_TRY(dalloc(&cells, g.top_cells * sizeof(uint32_t)));
    GC_TRY(dalloc(&request, g.mrpd_cells * sizeof(uint32_t)));
    GC_TRY(dalloc(&usage, (size_t)g.nslots * sizeof(uint32_t)));
    GC_TRY(dalloc(&order, (size_t)g.nslots * sizeof(uint32_t)));
    GC_TRY(dalloc(&order_tmp, (size_t)g.nslots * sizeof(uint32_t)));
    GC_TRY(dalloc(&meta, (size_t)g.nslots * sizeof(SlotMeta)));
    // +64 bytes: the ray caster reads texel pairs with 8-byte loads that may
    // run up to 5 bytes past the last row of the last slot
    GC_TRY(dalloc(&pool, (size_t)g.nslots * bvox * vc.encoded_bytes + 64));
    GC_TRY(dalloc(&d_geom, sizeof(gc_geometry)));
    // Surfaces and thrust vectors of the reference start zeroed
    // (Surface3dDevice.hpp:97-109, DynamicArray3dDevice).
    GC_HIP(hipMemsetAsync(mrpd, 0, g.top_cells * sizeof(uint4), stream));
    GC_HIP(hipMemsetAsync(cells, 0, g.top_cells * sizeof(uint32_t), stream));
    GC_TRY(dalloc(&d_unref, 16));
    GC_HIP(hipMemsetAsync(d_unref, 0, 16, stream));
    uint32_t lru_max = g.nslots;
    for (uint32_t i = 0; i < g.nb_tables; ++i) {
        // a table cache: zeroed entries and usage, LRU in block order with
        // no owner (CacheDevice.hpp:145-165)
        const gc_table_geometry &t = g.table[i];
        const size_t ne = (size_t)t.nblocks * table_b3(t);
        GC_TRY(dalloc(&tabs.words[i], ne * 4));
        GC_TRY(dalloc(&tabs.usage[i], (size_t)t.nblocks * 4));
        GC_TRY(dalloc(&tabs.order[i], (size_t)t.nblocks * 4));
        GC_TRY(dalloc(&tab_order_tmp[i], (size_t)t.nblocks * 4));
        GC_TRY(dalloc(&tabs.meta[i], (size_t)t.nblocks * sizeof(SlotMeta)));
        GC_HIP(hipMemsetAsync(tabs.words[i], 0, ne * 4, stream));
        GC_HIP(hipMemsetAsync(tabs.usage[i], 0, (size_t)t.nblocks * 4, stream));
        std::vector<uint32_t> ord(t.nblocks);
        std::vector<SlotMeta> mt(t.nblocks);
        for (uint32_t b = 0; b < t.nblocks; ++b) {
            ord[b] = b;
            mt[b].vol[0] = mt[b].vol[1] = mt[b].vol[2] = -1.f;
            mt[b].level = 0;
        }
        GC_HIP(hipMemcpyAsync(tabs.order[i], ord.data(), ord.size() * 4, hipMemcpyHostToDevice, stream));
        GC_HIP(hipMemcpyAsync(tabs.meta[i], mt.data(), mt.size() * sizeof(SlotMeta), hipMemcpyHostToDevice, stream));
        GC_HIP(hipStreamSynchronize(stream));
        lru_max = std::max(lru_max, t.nblocks);
    }
    GC_HIP(hipMemsetAsync(request, 0, g.mrpd_cells * sizeof(uint32_t), stream));
    GC_HIP(hipMemsetAsync(usage, 0, (size_t)g.nslots * sizeof(uint32_t), stream));
    GC_HIP(hipMemsetAsync(pool, 0, (size_t)g.nslots * bvox * vc.encoded_bytes + 64, stream));
    {
        // LRU initialisation, x fastest (CacheDevice.hpp:151-165)
        std::vector<uint32_t> ord(g.nslots);
        std::vector<SlotMeta> mt(g.nslots);
        for (uint32_t s = 0; s < g.nslots; ++s) {
            ord[s] = s;
            mt[s].vol[0] = mt[s].vol[1] = mt[s].vol[2] = -1.f;
            mt[s].level = 0;
        }
        GC_HIP(hipMemcpyAsync(order, ord.data(), ord.size() * 4, hipMemcpyHostToDevice, stream));
        GC_HIP(hipMemcpyAsync(meta, mt.data(), mt.size() * sizeof(SlotMeta), hipMemcpyHostToDevice, stream));
        GC_HIP(hipMemcpyAsync(d_geom, &g, sizeof(g), hipMemcpyHostToDevice, stream));
        GC_HIP(hipStreamSynchronize(stream));
    }
    lru_s.nblocks = (lru_max + 1023) / 1024;
    GC_TRY(dalloc(&lru_s.masks, ((size_t)lru_max + 63) / 64 * 8 + 8));
    GC_TRY(dalloc(&lru_s.counts, (size_t)lru_s.nblocks * 4 + 4));
    GC_TRY(dalloc(&lru_s.total, 16));
    miss_s.nblocks = (uint32_t)((g.mrpd_cells + 1023) / 1024);
    GC_TRY(dalloc(&miss_s.masks, ((size_t)g.mrpd_cells + 63) / 64 * 8 + 8));
    GC_TRY(dalloc(&miss_s.counts, (size_t)miss_s.nblocks * 4 + 4));
    GC_TRY(dalloc(&miss_s.total, 16));
    GC_TRY(dalloc(&d_misses, (size_t)cap * 16));
    GC_TRY(dalloc(&d_count, 16));
    GC_TRY(halloc(&h_misses, 16 + (size_t)cap * 16));
    // (sources that write bricks on the device fill the device staging
    // buffer directly: no host staging)
    if (!src->on_device()) GC_TRY(halloc(&h_staging, (size_t)cap * src->bb_bytes));
    GC_TRY(dalloc(&d_staging, (size_t)cap * src->bb_bytes));
    GC_TRY(halloc(&h_meta, (size_t)cap * 9 * 4));
    GC_TRY(dalloc(&d_meta, (size_t)cap * 9 * 4));
    if (g.nb_tables == 1) {
        GC_TRY(dalloc(&refs.cell, (size_t)cap * 3 * 4));
        refs.loc = refs.cell + cap;
        refs.rep = refs.cell + 2 * (size_t)cap;
        GC_TRY(dalloc(&refs.w0, (size_t)cap * 3 * 4));
        refs.rank = refs.w0 + cap;
        refs.ev = refs.w0 + 2 * (size_t)cap;
        GC_TRY(dalloc(&refs.inv, (size_t)g.table[0].nblocks * 4));
        GC_TRY(dalloc(&refs.flag, 16));
        GC_TRY(halloc(&h_ref, (size_t)cap * 3 * 4));
    }
    handler.c = this;
    if (!cc.sync_requests) handler.start();
    return GC_OK;
}

int Cache::lru_update()
{
    // CacheManager.hpp:387-406
    GC_HIP(launch_lru_update(order, order_tmp, usage, g.nslots, ts, lru_s, stream));
    std::swap(order, order_tmp);
    // the table caches after the data cache (the used-slot total the
    // counters read stays the data cache's: read it before)
    if (g.nb_tables) GC_HIP(hipMemcpyAsync(d_count + 1, lru_s.total, 4, hipMemcpyDeviceToDevice, stream));
    for (uint32_t i = 0; i < g.nb_tables; ++i) {
        GC_HIP(launch_lru_update(tabs.order[i], tab_order_tmp[i], tabs.usage[i], g.table[i].nblocks, ts, lru_s, stream));
        std::swap(tabs.order[i], tab_order_tmp[i]);
    }
    return GC_OK;
}

int Cache::prepare_bricks(bool &done)
{
    // prepare_bricks + create_request_list (CacheManager.hpp:410-491)
    GC_HIP(launch_miss_compact(request, g.mrpd_cells, ts, d_geom, limit, d_misses, miss_s, stream));
    GC_HIP(hipMemcpyAsync(h_misses, miss_s.total, 4, hipMemcpyDeviceToHost, stream));
    GC_HIP(hipMemcpyAsync(h_misses + 1, g.nb_tables ? d_count + 1 : lru_s.total, 4, hipMemcpyDeviceToHost,
                          stream));   // slots used
    GC_HIP(hipMemcpyAsync(h_misses + 4, d_misses, (size_t)cap * 16, hipMemcpyDeviceToHost, stream));
    GC_HIP(hipStreamSynchronize(stream));
    const uint32_t n = std::min(h_misses[0], limit);
    if (n == 0) {
        done = true;
        return GC_OK;
    }
    done = false;
    ++counters.requests;
    counters.last_used = h_misses[1];
    counters.last_requested = n;
    std::vector<uint32_t> list(h_misses + 4, h_misses + 4 + 4 * (size_t)n);
    if (cc.sync_requests) {
        GC_TRY(handler.handle(list));
        request_done = true;
    } else {
        handler.notify(std::move(list));
    }
    return GC_OK;
}

int Cache::insert_serial_needed(const uint32_t *asked, uint32_t n, int &serial) const
{
    // The parallel insert is exact when no two asked bricks share a page-table
    // cell under the FP32 position formula; otherwise replay the reference's
    // serial order.  The check runs between a launch and the next one: a
    // bitset over the cells (cleared again afterwards) keeps it to a few ns
    // per brick, a hash set for directories above 2^30 cells.
    serial = 0;
    if (g.mrpd_cells > (1ull << 30)) {
        std::unordered_set<uint64_t> seen;
        seen.reserve(n * 2);
        for (uint32_t k = 0; k < n; ++k) {
            float vol[3];
            volume_position(g, asked[4 * k + 3], asked + 4 * k, vol);
            if (!seen.insert(mrpd_index(g, asked[4 * k + 3], vol)).second) { serial = 1; break; }
        }
        return GC_OK;
    }
    const size_t words = (size_t)((g.mrpd_cells + 63) / 64);
    if (cell_bits.size() != words) cell_bits.assign(words, 0ull);
    cells_set.clear();
    for (uint32_t k = 0; k < n; ++k) {
        float vol[3];
        volume_position(g, asked[4 * k + 3], asked + 4 * k, vol);
        const uint64_t c = mrpd_index(g, asked[4 * k + 3], vol);
        if (c >= g.mrpd_cells) { serial = 1; break; }   // (never: positions lie in [0, 1))
        const uint64_t bit = 1ull << (c & 63);
        if (cell_bits[c >> 6] & bit) { serial = 1; break; }
        cell_bits[c >> 6] |= bit;
        cells_set.push_back(c);
    }
    for (const uint64_t c : cells_set) cell_bits[c >> 6] &= ~(1ull << (c & 63));
    return GC_OK;
}

int Cache::reference_plan(const uint32_t *asked, uint32_t n, bool &parallel)
{
    // Per brick the MRPD cell, the entry inside its table block and the first
    // brick of the load in the same cell, for the parallel referencing; two
    // bricks on one entry (FP32 positions) take the serial one.
    parallel = false;
    std::unordered_map<uint32_t, uint32_t> first;
    std::unordered_set<uint64_t> entries;
    first.reserve(n * 2);
    entries.reserve(n * 2);
    for (uint32_t k = 0; k < n; ++k) {
        float vol[3];
        volume_position(g, asked[4 * k + 3], asked + 4 * k, vol);
        const uint32_t c = top_index(g, asked[4 * k + 3], vol);
        const uint32_t l = table_local(g.table[0], asked[4 * k + 3], vol);
        if (c >= g.top_cells || l >= table_b3(g.table[0])) return GC_OK;   // (never: positions lie in [0, 1))
        if (!entries.insert((uint64_t)c << 32 | l).second) return GC_OK;
        h_ref[k] = c;
        h_ref[cap + k] = l;
        h_ref[2 * (size_t)cap + k] = first.emplace(c, k).first->second;
    }
    // (h_ref is rewritten only at the next load, after a synchronised update)
    for (int a = 0; a < 3; ++a)
        GC_HIP(hipMemcpyAsync(refs.cell + (size_t)a * cap, h_ref + (size_t)a * cap, (size_t)n * 4,
                              hipMemcpyHostToDevice, stream));
    parallel = true;
    return GC_OK;
}

int Cache::load_new_bricks()
{
    // CacheManager.hpp:503-632
    const uint32_t n = handler.n, nne = handler.nne;
    if (n == 0) return GC_OK;
    ++counters.loads;
    counters.bricks_loaded += nne;
    GC_HIP(hipStreamWaitEvent(stream, uploaded, 0));
    const uint32_t *d_asked = d_meta;
    const uint32_t *d_bbi = d_meta + 4 * (size_t)cap;
    const uint32_t *d_bbc = d_meta + 5 * (size_t)cap;
    if (g.nb_tables) {
        GC_HIP(launch_deref_tables(cells, tabs, order, g.nslots, nne, meta, d_geom, ts, stream));
        GC_HIP(launch_assign_copy(order, g.nslots, nne, d_asked, d_bbi, d_bbc, meta, pool, d_staging, d_geom,
                                  vc.big_brick_size, vc.encoded_bytes, stream));
        bool parallel = false;
        if (g.nb_tables == 1) GC_TRY(reference_plan(h_meta, n, parallel));
        if (parallel)
            GC_HIP(launch_reference_table1(mrpd, cells, tabs, order, g.nslots, n, nne, d_asked, meta, d_geom, ts,
                                           d_unref, refs, stream, g.table[0].nblocks));
        else
            GC_HIP(launch_reference_tables(mrpd, cells, tabs, order, g.nslots, n, nne, d_asked, meta, d_geom, ts,
                                           d_unref, stream));
    } else {
        GC_HIP(launch_deref(mrpd, cells, order, g.nslots, nne, meta, d_geom, stream));
        GC_HIP(launch_assign_copy(order, g.nslots, nne, d_asked, d_bbi, d_bbc, meta, pool, d_staging, d_geom,
                                  vc.big_brick_size, vc.encoded_bytes, stream));
        int serial = 0;
        GC_TRY(insert_serial_needed(h_meta, n, serial));
        GC_HIP(launch_insert(mrpd, cells, order, g.nslots, n, nne, d_asked, d_geom, serial, stream));
    }
    GC_HIP(hipEventRecord(consumed, stream));
    return GC_OK;
}

int Cache::update(int *done)
{
    // CacheManager.hpp:319-354
    ++counters.updates;
    GC_TRY(lru_update());
    if (handler.new_request.load()) {
        ++ts;
        if (done) *done = 0;
        return GC_OK;
    }
    if (handler.status != GC_OK) {
        const int rc = handler.status;
        handler.status = GC_OK;
        return fail(rc, handler.error);
    }
    bool check = false;
    bool rd;
    {
        std::lock_guard<std::mutex> lk(done_m);
        rd = request_done;
    }
    if (rd) {
        GC_TRY(load_new_bricks());
        std::lock_guard<std::mutex> lk(done_m);
        request_done = false;
        check = false;
    } else {
        GC_TRY(prepare_bricks(check));
    }
    ++ts;
    if (done) *done = check ? 1 : 0;
    return GC_OK;
}

// ------------------------------------------------------------------ handler
void Handler::start()
{
    run = true;
    new_request = false;
    worker = std::thread(&Handler::loop, this);
}

void Handler::stop()
{
    if (worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(m);
            incoming.clear();
            run = false;
            cv.notify_one();
        }
        worker.join();
    }
}

void Handler::notify(std::vector<uint32_t> &&list)
{
    std::lock_guard<std::mutex> lk(m);
    incoming = std::move(list);
    new_request = true;
    cv.notify_one();
}

void Handler::loop()
{
    hipSetDevice(c->device);
    for (;;) {
        std::vector<uint32_t> list;
        {
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return !run || new_request.load(); });
            if (!run) return;
            list = incoming;
        }
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = handle(list);
        c->fetch_ns += (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                           std::chrono::steady_clock::now() - t0).count();
        if (rc != GC_OK) {
            status = rc;
            error = gc_last_error();
        }
        {
            std::lock_guard<std::mutex> lk(c->done_m);
            c->request_done = rc == GC_OK;
        }
        {
            std::lock_guard<std::mutex> lk(m);
            new_request = false;
        }
        idle_cv.notify_all();
    }
}

int Handler::handle(const std::vector<uint32_t> &list)
{
    // RequestHandler::handle_request (RequestHandler.hpp:279-372)
    Source &S = *c->src;
    const uint32_t nreq = (uint32_t)(list.size() / 4);
    const uint32_t cap = c->cap;
    // the previous upload must be finished before pinned memory is rewritten
    if (upload_pending) {
        GC_HIP(hipEventSynchronize(c->uploaded));
        upload_pending = false;
    }
    // unique big bricks, first-occurrence order (std::map of :289).  The
    // list holds each requested cell once (a compaction of the request
    // buffer), so with big bricks of one brick every request is its own big
    // brick; otherwise a hash of the packed key finds the repeats.  This runs
    // serially between a launch and the next one.
    std::vector<std::array<uint32_t, 4>> uniq;
    std::vector<uint32_t> uid(nreq);
    uniq.reserve(nreq);
    if (S.BB[0] == 1 && S.BB[1] == 1 && S.BB[2] == 1) {
        for (uint32_t i = 0; i < nreq; ++i) {
            const uint32_t *a = &list[4 * i];
            uniq.push_back({a[0], a[1], a[2], a[3]});
            uid[i] = i;
        }
    } else {
        std::unordered_map<uint64_t, uint32_t> packed;
        std::map<std::array<uint32_t, 4>, uint32_t> wide;   // keys with a field of 2^16 or more
        packed.reserve((size_t)nreq * 2);
        for (uint32_t i = 0; i < nreq; ++i) {
            const uint32_t *a = &list[4 * i];
            const std::array<uint32_t, 4> key{a[0] / S.BB[0], a[1] / S.BB[1], a[2] / S.BB[2], a[3]};
            const uint32_t next = (uint32_t)uniq.size();
            uint32_t id;
            if (key[0] < 65536u && key[1] < 65536u && key[2] < 65536u && key[3] < 65536u)
                id = packed.emplace((uint64_t)key[0] | (uint64_t)key[1] << 16 | (uint64_t)key[2] << 32 |
                                        (uint64_t)key[3] << 48,
                                    next)
                         .first->second;
            else
                id = wide.emplace(key, next).first->second;
            if (id == next) uniq.push_back(key);
            uid[i] = id;
        }
    }
    const uint32_t nu = (uint32_t)uniq.size();
    std::vector<int> st(nu, GC_BRICK_SUCCESS);
    const float thr = c->cc.empty_threshold;
    auto fetch = [&](uint32_t u) {
        uint8_t *dst = c->h_staging + (uint64_t)u * S.bb_bytes;
        const uint32_t bbc[3] = {uniq[u][0], uniq[u][1], uniq[u][2]};
        int s = S.get(uniq[u][3], bbc, dst);
        if (s == GC_BRICK_SUCCESS) {
            // empty classification (BrickProcessorPredicate.hpp:11-45)
            bool all_below = true;
            for (uint64_t v = 0; v < S.bb_vox && all_below; ++v) {
                float x;
                if (S.bytes == 1) x = (float)dst[v];
                else if (S.bytes == 2) x = (float)reinterpret_cast<const uint16_t *>(dst)[v];
                else x = reinterpret_cast<const float *>(dst)[v];
                if (!(x < thr)) all_below = false;
            }
            if (all_below) s = GC_BRICK_EMPTY;
        }
        st[u] = s;
    };
    const uint32_t nth = S.thread_safe() ? std::max(1u, S.threads()) : 1u;
    // The device staging buffer is rewritten only after the last load
    // consumed it.
    GC_HIP(hipStreamWaitEvent(c->upload, c->consumed, 0));
    if (S.on_device()) {
        // the source writes the bricks straight into the staging buffer
        std::vector<uint32_t> keys(4 * (size_t)nu);
        for (uint32_t u = 0; u < nu; ++u) std::memcpy(&keys[4 * (size_t)u], uniq[u].data(), 16);
        std::vector<uint8_t> below(nu);
        GC_TRY(S.get_device(keys.data(), nu, c->d_staging, thr, below.data(), c->upload));
        for (uint32_t u = 0; u < nu; ++u) st[u] = below[u] ? GC_BRICK_EMPTY : GC_BRICK_SUCCESS;
        GC_HIP(hipEventRecord(c->uploaded, c->upload));
        upload_pending = true;
    } else {
        // Bricks go up on the side stream in chunks as soon as a chunk is
        // fetched, so the upload overlaps the fetches of the later chunks.
        const uint32_t chunk = 64;
        std::atomic<int> up_err{(int)hipSuccess};
        parallel_for_chunks(nu, nth, chunk, fetch, [&](uint32_t k) {
            const uint64_t u0 = (uint64_t)k * chunk, nb = std::min<uint64_t>(chunk, nu - u0);
            hipSetDevice(c->device);
            const hipError_t e = hipMemcpyAsync(c->d_staging + u0 * S.bb_bytes, c->h_staging + u0 * S.bb_bytes,
                                                nb * S.bb_bytes, hipMemcpyHostToDevice, c->upload);
            if (e != hipSuccess) up_err = (int)e;
        });
        GC_HIP(hipEventRecord(c->uploaded, c->upload));
        upload_pending = true;
        if (up_err != (int)hipSuccess) GC_HIP((hipError_t)up_err.load());
    }
    for (uint32_t u = 0; u < nu; ++u)
        if (st[u] == GC_BRICK_ERROR)
            return fail(GC_ERR_SOURCE, "brick source failed for level " + std::to_string(uniq[u][3]));
    // non-empty first, stable (RequestHandler.hpp:354-369)
    uint32_t *asked = c->h_meta;
    uint32_t *bbi = c->h_meta + 4 * (size_t)cap;
    uint32_t *bbc = c->h_meta + 5 * (size_t)cap;
    uint32_t k = 0;
    for (uint32_t i = 0; i < nreq; ++i)
        if (st[uid[i]] != GC_BRICK_EMPTY) {
            std::memcpy(asked + 4 * k, &list[4 * i], 16);
            bbi[k] = uid[i];
            ++k;
        }
    const uint32_t nne_ = k;
    for (uint32_t i = 0; i < nreq; ++i)
        if (st[uid[i]] == GC_BRICK_EMPTY) std::memcpy(asked + 4 * k++, &list[4 * i], 16);
    for (uint32_t u = 0; u < nu; ++u) std::memcpy(bbc + 4 * u, uniq[u].data(), 16);
    GC_HIP(hipMemcpyAsync(c->d_meta, c->h_meta, (size_t)cap * 9 * 4, hipMemcpyHostToDevice, c->upload));
    GC_HIP(hipEventRecord(c->uploaded, c->upload));
    upload_pending = true;
    n = nreq;
    nne = nne_;
    nuniq = nu;
    return GC_OK;
}

// ==================================================================== frame
constexpr uint32_t HAND_CTL_WORDS = 65 * 32;   // see Frame::hand_ctl
constexpr uint32_t LANES_PER_WG = 256;          // the ray casters' workgroups

struct Frame {
    uint32_t w = 0, h = 0, precision = 32;
    uint32_t *pixels = nullptr;
    float *tbuf = nullptr;
    void *cstate = nullptr;
    float4 *rpos = nullptr;                    // resume positions (views with resume = 1)
    double *qbuf = nullptr;                    // FP32: exact opacity products
    // launch scratch: [0..7] per-XCD packet cursors, [8], [9] ray-list counts
    uint32_t *ctl = nullptr;
    // hand-over scratch: [0] rays handed over in the launch, [1] their
    // cursor, [2] waves of the launch that ran out of rays; go flags at
    // 32 (k + 1), k < 64
    uint32_t *hand_ctl = nullptr;
    uint32_t *lists[2] = {nullptr, nullptr};   // rays stopped at a miss
    uint32_t *handed = nullptr;                // rays the lane march handed over in the launch
    float4 *handed_pos = nullptr;              // and their exact positions
    uint8_t *stop = nullptr;                   // per pixel: stopped at a miss in the last launch
    size_t cbytes() const { return (size_t)w * h * (precision == 64 ? 32 : 16); }
    ~Frame()
    {
        hipFree(pixels);
        hipFree(tbuf);
        hipFree(cstate);
        hipFree(rpos);
        hipFree(qbuf);
        hipFree(ctl);
        hipFree(lists[0]);
        hipFree(lists[1]);
        hipFree(handed);
        hipFree(handed_pos);
        hipFree(hand_ctl);
        hipFree(stop);
    }
};

// Hand-over of a lane-per-ray launch L (grid workgroups of 4 waves) to the
// group march H that continues its handed-over rays: L gets the budget, the
// trigger and the hand-over buffers (list and positions of capacity >= the
// launch's rays, hand_ctl as Frame::hand_ctl, cleared before L runs), H the
// same launch over the handed-over list.
// Whether lane-per-ray launch L over cache C can hand rays over: u8/u16
// voxels with covering >= 1, the pipelined march or the walk through table
// caches, and a group march for the continuation.
static bool handoff_possible(const Cache &C, const DvrLaunch &L)
{
    return C.handoff_after != 0 && (L.pipelined || L.pt.n != 0) && C.vc.encoded_bytes != 4 && C.g.covering >= 1 &&
           dvr_ray_supported(L, C.vc.encoded_bytes);
}

// Lanes per ray of the continuation (the table-walk group marches have 16 or
// more).
static uint32_t handoff_group_of(const Cache &C, const DvrLaunch &L)
{
    return L.pt.n != 0 ? std::max(16u, C.handoff_group) : C.handoff_group;
}

static void handoff_launches(const Cache &C, DvrLaunch &L, DvrLaunch &H, uint32_t grid, uint32_t *handed,
                             float4 *handed_pos, uint32_t *hand_ctl)
{
    L.yield_after = C.handoff_after;
    L.yield_exited = C.handoff_exited_pct
                         ? std::max(1u, (uint32_t)((uint64_t)grid * 4u * C.handoff_exited_pct / 100u))
                         : 0u;
    L.exited = hand_ctl + 2;
    L.go = hand_ctl + 32;
    L.yield_list = handed;
    L.yield_count = hand_ctl;
    L.yield_pos = handed_pos;
    H = L;
    H.yield_after = H.yield_exited = 0;
    H.exited = H.go = nullptr;
    H.yield_list = H.yield_count = nullptr;
    H.list = handed;
    H.list_count = hand_ctl;
    H.queue = hand_ctl + 1;
    H.n_packets = 0;
    H.continuation = 1;
}

// Exact opacity threshold of a view for FP32 compositing: a ray continues
// while its opacity product Q > (1 - alpha_stop)^(1/k0) (see kernels_dvr.hip,
// opacity_factor); an alpha stop of 1 continues while Q > 0, above 1 always.
double opacity_threshold(float alpha_stop, float k0)
{
    const double a = (double)alpha_stop;
    if (a > 1.0) return -1.0;
    if (a == 1.0) return 0.0;
    return std::pow(1.0 - a, 1.0 / (double)k0);
}

// Exact opacity decisions hold when every level's exponent is a power-of-two
// multiple of level 0's and every threshold stays a normal double.
void finish_exact(DvrLaunch &L, const DvrView *views, uint32_t n)
{
    for (uint32_t i = 0; i < n; ++i)
        if (views[i].width && views[i].thr >= 0.0 && views[i].thr < 1e-290) L.exact_alpha = 0;
}

// Shared part of a ray-caster launch over cache C: device state, timestamp
// and the LOD tables of VolumeRayCaster.cu:136-161 (start-up box [-1,1]^3).
static void launch_base(const Cache &C, DvrLaunch &L)
{
    L.mrpd = C.mrpd;
    L.cells = C.cells;
    L.brick_size = C.g.brick_size;
    L.covering = C.g.covering;
    L.request = C.request;
    L.usage = C.usage;
    L.pool = C.pool;
    L.geom = C.d_geom;
    L.ts = C.ts;
    const float smin = 2.f;
    L.tstep = smin / 256.f;
    L.nb_lods = C.vc.nb_levels;
    for (uint32_t i = 0; i < C.vc.nb_levels; ++i) {
        const uint32_t *ls = C.vc.initial_sizes[i];
        for (int a = 0; a < 3; ++a) L.inv_levels[i][a] = smin / (float)ls[a];
        const uint32_t m = std::min(ls[0], std::min(ls[1], ls[2]));
        L.lod_step[i] = (smin / (float)m) / 2.f;
        L.kexp[i] = L.lod_step[i] / L.tstep;
    }
    // exact opacity decisions: kexp[l] / kexp[0] = 2^esq[l] on every level
    L.exact_alpha = 1;
    for (uint32_t i = 0; i < C.vc.nb_levels; ++i) {
        const double r = (double)L.kexp[i] / (double)L.kexp[0];
        int e = 0;
        const double m = std::frexp(r, &e);   // r = m * 2^e, m in [0.5, 1)
        if (m != 0.5 || e < 1 || e > 31) {
            L.exact_alpha = 0;
            L.esq[i] = 0;
        } else {
            L.esq[i] = e - 1;
        }
    }
    L.qskip = std::pow(1.0 - (double)0.001f, 1.0 / (double)L.kexp[0]);
    // The pipelined march takes the request cell to be the page cell (true when
    // realDivVirtual is exactly 1, which the bricker's padded level sizes give)
    // and forms page indices with 24-bit multiplies.
    bool rdv_one = true;
    for (uint32_t l = 0; l < C.vc.nb_levels; ++l)
        for (int a = 0; a < 3; ++a) rdv_one = rdv_one && C.g.rdv[l][a] == 1.0f;
    L.pipelined = rdv_one && C.g.mrpd_dim[0] < (1u << 24) &&
                  (uint64_t)C.g.mrpd_dim[1] * C.g.mrpd_dim[2] < (1u << 24);
    // table caches: the walk through them (the non-pipelined march)
    L.pt = PageTables{};
    L.pt.n = C.g.nb_tables;
    L.pt.top_x = C.g.top_cells_dim[0];
    L.pt.top_y = C.g.top_cells_dim[1];
    for (uint32_t i = 0; i < C.g.nb_tables; ++i) {
        L.pt.pow2[i] = 1;
        for (int a = 0; a < 3; ++a) {
            L.pt.block[i][a] = C.g.table[i].block[a];
            if (L.pt.block[i][a] & (L.pt.block[i][a] - 1)) L.pt.pow2[i] = 0;
        }
        L.pt.b3[i] = table_b3(C.g.table[i]);
        L.pt.words[i] = C.tabs.words[i];
        L.pt.usage[i] = C.tabs.usage[i];
    }
    if (C.g.nb_tables) L.pipelined = 0;
}

int cache_launch_base(gc_cache *c, DvrLaunch &L, hipStream_t &stream, int &device, uint32_t &bytes);

static void view_of(const gc_view &v, const float *d_tf, uint32_t tf_size, const Frame &F, float k0, DvrView &out)
{
    out.width = v.width;
    out.height = v.height;
    const uint32_t tile = v.compositing == 1 ? 0u : v.tile;   // maximum projection: 8x8 tiles
    out.tiles_x = tile_cols(tile, v.width);
    out.tile_pair = tile_pair(tile);
    out.packet_off = 0;
    std::memcpy(out.m, v.inv_mv, sizeof(out.m));
    std::memcpy(out.bmin, v.bbox_min, sizeof(out.bmin));
    std::memcpy(out.bmax, v.bbox_max, sizeof(out.bmax));
    out.alpha_stop = v.alpha_stop;
    out.tf_size = tf_size;
    out.tf = reinterpret_cast<const float4 *>(d_tf);
    out.pixels = F.pixels;
    out.tbuf = F.tbuf;
    out.cstate = F.cstate;
    out.rpos = v.resume ? F.rpos : nullptr;
    out.qbuf = F.precision == 32 ? F.qbuf : nullptr;
    out.thr = opacity_threshold(v.alpha_stop, k0);
}

static uint32_t packets_of(const gc_view &v) { return tile_packets(v.tile, v.width, v.height); }



// ------------------------------------------------------------------ batches
// A set of view slots over one cache.  Each launch covers the rays the
// previous launch stopped at a miss (any slot) and the tiles of the slots
// that start in it.
struct Batch {
    Cache *C = nullptr;                 // borrowed: the handle's cache
    gc_cache *h = nullptr;
    uint32_t precision = 32, max_views = 0;
    uint64_t max_pixels = 0;
    uint32_t *pixels = nullptr;
    float *tbuf = nullptr;
    void *cstate = nullptr;
    float4 *rpos = nullptr;             // resume positions (views with resume = 1)
    double *qbuf = nullptr;             // FP32: exact opacity products
    uint32_t *lists[2] = {nullptr, nullptr};
    uint32_t *handed = nullptr;         // rays handed over in a launch (Frame::handed)
    float4 *handed_pos = nullptr;
    uint32_t *hand_ctl = nullptr;       // see Frame::hand_ctl
    uint8_t *stop = nullptr;            // per pixel: stopped at a miss in the last launch
    // [0..7] per-XCD packet cursors, [8], [9] ray-list counts, [10] timestamp,
    // [16..16+max_views) rays stopped per slot in the last launch
    uint32_t *ctl = nullptr;
    uint32_t *h_stopped = nullptr;      // pinned copy of the per-slot counters
    DvrView *d_views = nullptr, *h_views = nullptr;   // h_views pinned
    gc_render_stats *d_stats = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, views_free = nullptr, launched = nullptr;
    struct Slot {
        bool used = false, starting = false;
        DvrView v{};
        uint64_t off = 0, npix = 0;
    };
    std::vector<Slot> slots;
    uint32_t table = 0;                 // slots in the device table (highest used + 1)
    uint32_t tile = 0, march = 8, refill_min = 16;
    int cur = -1;                       // list holding the last launch's stopped rays (-1: none)
    bool use_graph = false;
    std::map<std::array<uint32_t, 11>, hipGraphExec_t> graphs;
    size_t cbytes_per_pixel() const { return precision == 64 ? 32 : 16; }
    uint64_t slot_pixels() const { return max_pixels / max_views; }
    void drop_graphs()
    {
        for (auto &g : graphs) hipGraphExecDestroy(g.second);
        graphs.clear();
    }
    ~Batch()
    {
        drop_graphs();
        hipFree(pixels);
        hipFree(tbuf);
        hipFree(cstate);
        hipFree(rpos);
        hipFree(qbuf);
        hipFree(lists[0]);
        hipFree(lists[1]);
        hipFree(handed);
        hipFree(handed_pos);
        hipFree(hand_ctl);
        hipFree(stop);
        hipFree(ctl);
        hipFree(d_views);
        hipFree(d_stats);
        if (h_views) hipHostFree(h_views);
        if (h_stopped) hipHostFree(h_stopped);
        if (ev0) hipEventDestroy(ev0);
        if (ev1) hipEventDestroy(ev1);
        if (views_free) hipEventDestroy(views_free);
        if (launched) hipEventDestroy(launched);
    }
};

static int batch_place(Batch &B, uint32_t slot, const gc_view &v, const float *d_tf, uint32_t tf_size, uint64_t off)
{
    if (v.width == 0 || v.height == 0 || (uint64_t)v.width * v.height >= (1ull << DVR_PIXEL_BITS))
        return fail(GC_ERR_ARG, "slot " + std::to_string(slot) + ": bad view size");
    if (v.precision != B.precision) return fail(GC_ERR_ARG, "slot " + std::to_string(slot) + ": precision differs from the batch");
    if (!d_tf || tf_size == 0) return fail(GC_ERR_ARG, "slot " + std::to_string(slot) + ": no transfer function");
    if (v.compositing != 0)
        return fail(GC_ERR_ARG, "slot " + std::to_string(slot) + ": batches render DVR only (compositing 0)");
    Batch::Slot &S = B.slots[slot];
    S.used = true;
    S.starting = true;
    S.off = off;
    S.npix = (uint64_t)v.width * v.height;
    DvrView &D = S.v;
    D.width = v.width;
    D.height = v.height;
    std::memcpy(D.m, v.inv_mv, sizeof(D.m));
    std::memcpy(D.bmin, v.bbox_min, sizeof(D.bmin));
    std::memcpy(D.bmax, v.bbox_max, sizeof(D.bmax));
    D.alpha_stop = v.alpha_stop;
    D.tf_size = tf_size;
    D.tf = reinterpret_cast<const float4 *>(d_tf);
    D.pixels = B.pixels + off;
    D.tbuf = B.tbuf + off;
    D.cstate = static_cast<uint8_t *>(B.cstate) + off * B.cbytes_per_pixel();
    D.rpos = v.resume ? B.rpos + off : nullptr;
    D.qbuf = B.precision == 32 ? B.qbuf + off : nullptr;
    D.stop = B.stop + off;
    {
        DvrLaunch K{};
        launch_base(*B.C, K);
        D.thr = opacity_threshold(v.alpha_stop, K.kexp[0]);
    }
    B.table = std::max(B.table, slot + 1);
    // a starting view begins from an empty progressive state
    hipStream_t s = B.C->stream;
    GC_HIP(launch_fill32(reinterpret_cast<uint32_t *>(D.tbuf), (uint32_t)S.npix, 0, s));
    GC_HIP(launch_fill32(static_cast<uint32_t *>(D.cstate), (uint32_t)(S.npix * B.cbytes_per_pixel() / 4), 0, s));
    return GC_OK;
}

// One launch: the previous launch's stopped rays, then the starting slots'
// tiles.  No cache update.
static int batch_launch(Batch &B, gc_frame_report *rep, bool skip_list = false)
{
    Cache &C = *B.C;
    hipStream_t s = C.stream;
    // view table with this launch's tile ranges
    GC_HIP(hipEventSynchronize(B.views_free));
    uint32_t packets = 0, stop_tiles_n = 0;
    for (uint32_t i = 0; i < B.table; ++i) {
        Batch::Slot &S = B.slots[i];
        DvrView &D = B.h_views[i];
        D = S.v;
        D.packet_off = packets;
        D.stop_tile_off = stop_tiles_n;
        if (!S.used) {
            D.width = D.height = D.tiles_x = 0;
            D.stop = nullptr;
            continue;
        }
        stop_tiles_n += stop_tiles(D.width, D.height);
        D.tiles_x = tile_cols(B.tile, D.width);
        D.tile_pair = tile_pair(B.tile);
        if (S.starting) packets += tile_packets(B.tile, D.width, D.height);
    }
    // skip_list: leave the waiting rays in their list (their bricks are not
    // loaded yet); this launch's stopped rays are appended to it
    const bool has_list = B.cur >= 0 && !skip_list;
    const bool append = B.cur >= 0 && skip_list;
    if (packets == 0 && !has_list) return GC_OK;
    uint32_t tf_tables = 0;
    const uint32_t tf_entries = dvr_assign_tf_lds(B.h_views, B.table, C.vc.encoded_bytes, &tf_tables);
    GC_HIP(hipMemcpyAsync(B.d_views, B.h_views, (size_t)B.table * sizeof(DvrView), hipMemcpyHostToDevice, s));
    GC_HIP(hipEventRecord(B.views_free, s));
    DvrLaunch L{};
    launch_base(C, L);
    L.tf_lds_entries = tf_entries;
    L.tf_lds_tables = tf_tables;
    L.views = B.d_views;
    L.n_views = B.table;
    L.n_packets = packets;
    L.stats = B.d_stats;
    L.queue = B.ctl;
    L.refill_min = B.refill_min;
    L.march = B.march;
    for (uint32_t i = 0; i < B.table; ++i)
        if (B.slots[i].used && B.slots[i].v.rpos) L.exact_resume = 1;
    finish_exact(L, B.h_views, B.table);
    const int out = has_list ? 1 - B.cur : (append ? B.cur : 0);
    if (has_list) {
        L.list = B.lists[B.cur];
        L.list_count = B.ctl + 8 + B.cur;
    }
    // the stopped rays, in screen order, form the next launch's list (appended
    // to the waiting rays when the list was skipped)
    CollectArgs ca{};
    ca.views = B.d_views;
    ca.n_views = B.table;
    ca.n_tiles = stop_tiles_n;
    ca.out_list = B.lists[out];
    ca.out_count = B.ctl + 8 + out;
    ca.view_stopped = B.ctl + 16;
    uint64_t listed = 0;
    if (has_list)
        for (uint32_t i = 0; i < B.table; ++i)
            if (B.slots[i].used && !B.slots[i].starting) listed += B.slots[i].npix;
    // a launch that only resumes a short list of rays: one wave per ray
    uint32_t list_rays = 0;
    if (has_list && packets == 0) {
        GC_HIP(hipEventSynchronize(B.launched));
        for (uint32_t i = 0; i < B.table; ++i) list_rays += B.h_stopped[i];
    }
    // (a captured graph keeps 16 lanes per ray, whatever the list length)
    const uint32_t ray_group = B.use_graph ? 16u : 0u;
    const bool ray_kernel = has_list && packets == 0 &&
                            list_rays <= (B.use_graph ? std::min(C.ray_list_max, RAY_G16_MAX) : C.ray_list_max) &&
                            dvr_ray_supported(L, C.vc.encoded_bytes);
    const uint32_t lds = dvr_lds_bytes(L.nb_lods, B.table, tf_entries);
    // (a captured graph keeps one grid for every list length)
    const uint32_t grid = ray_kernel ? dvr_ray_grid(C.vc.encoded_bytes, B.precision, B.use_graph ? ~0u - 64u : list_rays,
                                                    dvr_ray_lds_bytes(B.table, tf_entries),
                                                    ray_group ? ray_group : C.group_for(list_rays, L.pt.n != 0))
                                     : dvr_grid(C.vc.encoded_bytes, B.precision, B.tile, packets + (uint32_t)((listed + 63) / 64), lds,
                                                tf_tables);
    // lane-per-ray launches hand their tail over to a group march, as the
    // frame loop's, when they resume rays or hold few rays per lane (a launch
    // of many fresh views, some 100 rays per lane, spends little time in its
    // tail and runs faster without the hand-over's registers)
    const bool handoff = !ray_kernel && handoff_possible(C, L) &&
                         (has_list || (uint64_t)packets * 64u <= (uint64_t)grid * LANES_PER_WG * 16u);
    DvrLaunch H{};
    uint32_t h_grid = 0;
    if (handoff) {
        handoff_launches(C, L, H, grid, B.handed, B.handed_pos, B.hand_ctl);
        h_grid = dvr_ray_grid(C.vc.encoded_bytes, B.precision, ~0u - 64u, dvr_ray_lds_bytes(B.table, tf_entries),
                              handoff_group_of(C, L));
    }
    auto launch_k = [&]() -> hipError_t {
        hipError_t e = hipSuccess;
        if (ray_kernel) {
            e = launch_dvr_ray(L, C.vc.encoded_bytes, B.precision, B.use_graph ? ~0u - 64u : list_rays, grid, s,
                               ray_group ? ray_group : C.group_for(list_rays, L.pt.n != 0));
        } else {
            if (handoff) e = launch_fill32(B.hand_ctl, HAND_CTL_WORDS, 0, s);
            if (e == hipSuccess) e = launch_dvr(L, C.vc.encoded_bytes, B.precision, B.tile, grid, s);
            if (e == hipSuccess && handoff)
                e = launch_dvr_ray(H, C.vc.encoded_bytes, B.precision, ~0u - 64u, h_grid, s, handoff_group_of(C, L));
        }
        return e == hipSuccess ? launch_collect_stopped(ca, s) : e;
    };
    GC_HIP(hipEventRecord(B.ev0, s));
    if (B.use_graph) {
        // the timestamp is read on the device so that a captured launch
        // stays valid across updates
        GC_HIP(launch_fill32(B.ctl + 10, 1, C.ts, s));
        L.ts_dev = B.ctl + 10;
        H.ts_dev = L.ts_dev;
        const std::array<uint32_t, 11> key{B.table, B.tile + (ray_kernel ? 8u : 0u), B.march, B.refill_min, packets, grid,
                                           has_list ? (uint32_t)B.cur : 2u, (uint32_t)out + (append ? 4u : 0u),
                                           tf_entries, stop_tiles_n,
                                           handoff ? (C.handoff_after << 16) ^ (C.handoff_group << 8) ^
                                                         C.handoff_exited_pct
                                                   : 0u};
        auto it = B.graphs.find(key);
        if (it == B.graphs.end()) {
            hipGraph_t g = nullptr;
            GC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            hipError_t e = launch_fill32(B.ctl, 8, 0, s);
            if (e == hipSuccess && !append) e = launch_fill32(B.ctl + 8 + out, 1, 0, s);
            if (e == hipSuccess && !append) e = launch_fill32(B.ctl + 16, B.max_views, 0, s);
            if (e == hipSuccess) e = launch_k();
            const hipError_t e_end = hipStreamEndCapture(s, &g);
            GC_HIP(e);
            GC_HIP(e_end);
            hipGraphExec_t x = nullptr;
            const hipError_t e_inst = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            GC_HIP(e_inst);
            it = B.graphs.emplace(key, x).first;
        }
        GC_HIP(hipGraphLaunch(it->second, s));
    } else {
        GC_HIP(launch_fill32(B.ctl, 8, 0, s));
        if (!append) {
            // the per-slot counters count the rays in the output list
            GC_HIP(launch_fill32(B.ctl + 8 + out, 1, 0, s));
            GC_HIP(launch_fill32(B.ctl + 16, B.max_views, 0, s));
        }
        GC_HIP(launch_k());
    }
    GC_HIP(hipEventRecord(B.ev1, s));
    GC_HIP(hipMemcpyAsync(B.h_stopped, B.ctl + 16, (size_t)B.max_views * 4, hipMemcpyDeviceToHost, s));
    GC_HIP(hipEventRecord(B.launched, s));
    B.cur = out;
    for (uint32_t i = 0; i < B.table; ++i) B.slots[i].starting = false;
    if (rep) ++rep->launches;
    return GC_OK;
}

// A slot whose rays stopped at a miss in the last launch is still in that
// launch's ray list and cannot be replaced.
static int slot_busy(Batch &B, uint32_t slot, bool &busy)
{
    busy = false;
    const Batch::Slot &S = B.slots[slot];
    if (!S.used || S.starting || B.cur < 0) return GC_OK;
    GC_HIP(hipEventSynchronize(B.launched));
    busy = B.h_stopped[slot] != 0;
    return GC_OK;
}

// Kernel time of the last launch; waits for it.
static int batch_kernel_ms(Batch &B, double &ms)
{
    GC_HIP(hipEventSynchronize(B.ev1));
    float f = 0.f;
    GC_HIP(hipEventElapsedTime(&f, B.ev0, B.ev1));
    ms = f;
    return GC_OK;
}

} // namespace gc

struct gc_cache {
    gc::Cache impl;
};

namespace gc {
int cache_launch_base(gc_cache *c, DvrLaunch &L, hipStream_t &stream, int &device, uint32_t &bytes)
{
    if (!c) return fail(GC_ERR_ARG, "null cache");
    Cache &C = c->impl;
    launch_base(C, L);
    stream = C.stream;
    device = C.device;
    bytes = C.vc.encoded_bytes;
    return GC_OK;
}
} // namespace gc
struct gc_frame {
    gc::Frame impl;
};
struct gc_batch {
    gc::Batch impl;
};

using namespace gc;

// ===================================================================== C ABI
extern "C" {

const char *gc_last_error(void) { return g_error.c_str(); }
const char *gc_version(void) { return "gcvol 0.1 (gfx950)"; }

int gc_volume_cfg_init(gc_volume_cfg *out, const uint32_t dims[3], uint32_t brick_size, uint32_t covering,
                       const uint32_t big_brick[3], uint32_t encoded_bytes)
{
    if (!out || !dims || !big_brick) return fail(GC_ERR_ARG, "null argument");
    if (brick_size <= 2 * covering) return fail(GC_ERR_CONFIG, "BrickSize must exceed 2*VoxelCovering");
    const uint32_t m = std::max(dims[0], std::max(dims[1], dims[2]));
    if (m / brick_size == 0) return fail(GC_ERR_CONFIG, "volume smaller than a brick");
    std::memset(out, 0, sizeof(*out));
    const uint32_t nl = level_count(dims, brick_size);
    if (nl > GC_MAX_LEVELS) return fail(GC_ERR_CONFIG, "too many levels");
    for (int a = 0; a < 3; ++a) {
        out->brick_size[a] = brick_size;
        out->big_brick_size[a] = big_brick[a];
        out->covering[a] = covering;
    }
    out->encoded_bytes = encoded_bytes;
    out->channels = 1;
    out->nb_levels = nl;
    level_sizes(dims, nl, out->initial_sizes);
    real_sizes(nl, out->initial_sizes, brick_size, covering, big_brick, out->real_sizes);
    return GC_OK;
}

int gc_pool_shape(uint64_t budget_bytes, uint64_t brick_bytes, uint32_t out[3])
{
    if (!out || brick_bytes == 0) return fail(GC_ERR_ARG, "bad argument");
    pool_shape(budget_bytes, brick_bytes, out);
    return GC_OK;
}

int gc_cfg_get(const char *cfg_path, const char *key, char *value, size_t value_size)
{
    if (!cfg_path || !key || !value || value_size == 0) return fail(GC_ERR_ARG, "bad argument");
    std::string v;
    if (!cfg_lookup(cfg_path, key, v)) return fail(GC_ERR_ARG, std::string("key not found: ") + key);
    std::snprintf(value, value_size, "%s", v.c_str());
    return GC_OK;
}

// Level records of a bricks.xml (<Levels><Level initial_size_X=".." ...
// real_size_Z=".."/>..</Levels>, VolumeConfiguration.cpp:64-98).
static bool read_bricks_xml(const std::string &path, gc_volume_cfg &vc)
{
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    std::fclose(f);
    auto attr = [](const std::string &tag, const char *name, uint32_t &out) {
        const std::string key = std::string(name) + "=\"";
        const size_t p = tag.find(key);
        if (p == std::string::npos) return false;
        out = (uint32_t)std::strtoul(tag.c_str() + p + key.size(), nullptr, 10);
        return true;
    };
    const size_t lv = text.find("<Levels");
    if (lv == std::string::npos) return false;
    const size_t lv_end = text.find("</Levels>", lv);
    uint32_t nl = 0;
    for (size_t p = text.find("<Level ", lv); p != std::string::npos && (lv_end == std::string::npos || p < lv_end);
         p = text.find("<Level ", p + 1)) {
        const std::string tag = text.substr(p, text.find('>', p) - p);
        if (nl >= GC_MAX_LEVELS) return false;
        const char *in[3] = {"initial_size_X", "initial_size_Y", "initial_size_Z"};
        const char *re[3] = {"real_size_X", "real_size_Y", "real_size_Z"};
        for (int a = 0; a < 3; ++a)
            if (!attr(tag, in[a], vc.initial_sizes[nl][a]) || !attr(tag, re[a], vc.real_sizes[nl][a])) return false;
        ++nl;
    }
    if (nl == 0) return false;
    vc.nb_levels = nl;
    return true;
}

int gc_volume_cfg_load(const char *cfg_path, gc_volume_cfg *out, char *volume_path, size_t volume_path_size)
{
    if (!cfg_path || !out) return fail(GC_ERR_ARG, "null argument");
    if (std::FILE *f = std::fopen(cfg_path, "rb")) std::fclose(f);
    else return fail(GC_ERR_CONFIG, std::string("VolumeConfiguration: unable to open ") + cfg_path);
    auto need = [&](const char *key, std::string &v) { return cfg_lookup(cfg_path, key, v); };
    auto need_u = [&](const char *key, uint32_t &v) {
        std::string s;
        if (!need(key, s)) return false;
        v = (uint32_t)std::strtoul(s.c_str(), nullptr, 10);
        return true;
    };
    uint32_t bs = 0, bb[3] = {0, 0, 0}, cov = 0, bytes = 0, channels = 0;
    std::string file, wd;
    const struct { const char *key; uint32_t *v; } req[] = {
        {"BrickSize", &bs}, {"BigBrickSizeX", &bb[0]}, {"BigBrickSizeY", &bb[1]}, {"BigBrickSizeZ", &bb[2]},
        {"VoxelCovering", &cov}, {"NumberEncodedBytes", &bytes}, {"NumberChannels", &channels}};
    for (const auto &r : req)
        if (!need_u(r.key, *r.v))
            return fail(GC_ERR_CONFIG, std::string("VolumeConfiguration: Missing ") + r.key + " in configuration.");
    if (!need("VolumeFile", file)) return fail(GC_ERR_CONFIG, "VolumeConfiguration: Missing VolumeFile in configuration.");
    need("WorkingDirectory", wd);
    if (bytes != 1 && bytes != 2 && bytes != 4) return fail(GC_ERR_CONFIG, "NumberEncodedBytes must be 1, 2 or 4");
    if (channels != 1) return fail(GC_ERR_CONFIG, "NumberChannels must be 1");
    if (bb[0] == 0 || bb[1] == 0 || bb[2] == 0) return fail(GC_ERR_CONFIG, "BigBrickSize must be positive");
    const size_t slash = file.find_last_of("/\\"), dot = file.find_last_of('.');
    const std::string base = file.substr(slash + 1, dot - (slash + 1));
    const std::string bricks_xml = wd + base + "/bricks_" + std::to_string(bs) + "_" + std::to_string(bs) + "_" +
                                   std::to_string(bs) + "/bricks.xml";
    gc_volume_cfg vc;
    std::memset(&vc, 0, sizeof(vc));
    if (read_bricks_xml(bricks_xml, vc)) {
        for (int a = 0; a < 3; ++a) {
            vc.brick_size[a] = bs;
            vc.big_brick_size[a] = bb[a];
            vc.covering[a] = cov;
        }
        vc.encoded_bytes = bytes;
        vc.channels = channels;
    } else {
        uint32_t dims[3];
        if (!need_u("size_X", dims[0]) || !need_u("size_Y", dims[1]) || !need_u("size_Z", dims[2]))
            return fail(GC_ERR_CONFIG, "VolumeConfiguration: no " + bricks_xml + " and no size_X/Y/Z in configuration.");
        const int rc = gc_volume_cfg_init(&vc, dims, bs, cov, bb, bytes);
        if (rc != GC_OK) return rc;
    }
    *out = vc;
    if (volume_path && volume_path_size) std::snprintf(volume_path, volume_path_size, "%s%s", wd.c_str(), file.c_str());
    return GC_OK;
}

int gc_device_count(int *count)
{
    if (!count) return fail(GC_ERR_ARG, "null argument");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return GC_OK;
}
int gc_set_device(int device) { GC_HIP(hipSetDevice(device)); return GC_OK; }
int gc_malloc(void **dptr, size_t bytes) { return dalloc(dptr, bytes); }
int gc_free(void *dptr) { GC_HIP(hipFree(dptr)); return GC_OK; }
int gc_memcpy_h2d(void *dst, const void *src, size_t bytes) { GC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); return GC_OK; }
int gc_memcpy_d2h(void *dst, const void *src, size_t bytes) { GC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); return GC_OK; }
int gc_memcpy_d2d(void *dst, const void *src, size_t bytes) { GC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice)); return GC_OK; }
int gc_memset(void *dptr, int value, size_t bytes) { GC_HIP(hipMemset(dptr, value, bytes)); return GC_OK; }
int gc_synchronize(void) { GC_HIP(hipDeviceSynchronize()); return GC_OK; }
int gc_stream_create(void **stream)
{
    if (!stream) return fail(GC_ERR_ARG, "null argument");
    hipStream_t s;
    GC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return GC_OK;
}
int gc_stream_destroy(void *stream) { GC_HIP(hipStreamDestroy((hipStream_t)stream)); return GC_OK; }
int gc_stream_synchronize(void *stream) { GC_HIP(hipStreamSynchronize((hipStream_t)stream)); return GC_OK; }

int gc_mip_reduce(const void *d_in, const uint32_t dims[3], uint32_t encoded_bytes, void *d_out, void *stream)
{
    if (!d_in || !d_out || !dims) return fail(GC_ERR_ARG, "null argument");
    GC_HIP(launch_mip(d_in, dims, encoded_bytes, d_out, (hipStream_t)stream));
    return GC_OK;
}

int gc_brick_cut(const void *d_level, const uint32_t dims[3], uint32_t encoded_bytes, uint32_t brick_size,
                 uint32_t covering, const uint32_t *d_bricks_xyz, uint32_t nb_bricks, void *d_out, void *stream)
{
    if (!d_level || !dims || !d_bricks_xyz || !d_out) return fail(GC_ERR_ARG, "null argument");
    if (brick_size <= 2 * covering) return fail(GC_ERR_CONFIG, "BrickSize must exceed 2*VoxelCovering");
    GC_HIP(launch_brick_cut(d_level, dims, encoded_bytes, brick_size, covering, d_bricks_xyz, nb_bricks, d_out,
                            (hipStream_t)stream));
    return GC_OK;
}

int gc_source_create_pyramid(const gc_volume_cfg *vc, const void *const *levels, gc_source **out)
{
    if (!vc || !levels || !out) return fail(GC_ERR_ARG, "null argument");
    auto s = std::make_unique<PyramidSource>();
    s->init(*vc);
    for (uint32_t l = 0; l < vc->nb_levels; ++l) {
        if (!levels[l]) return fail(GC_ERR_ARG, "null level pointer");
        s->levels.push_back(static_cast<const uint8_t *>(levels[l]));
    }
    *out = new gc_source{std::move(s)};
    return GC_OK;
}

int gc_source_create_synthetic(const gc_volume_cfg *vc, const gc_synth_params *p, gc_source **out)
{
    if (!vc || !p || !out) return fail(GC_ERR_ARG, "null argument");
    auto s = std::make_unique<SynthSource>();
    s->init(*vc);
    s->synth = std::make_unique<Synth>(*p, vc->initial_sizes[0], vc->encoded_bytes);
    s->nthreads = p->threads ? p->threads : hw_threads();
    *out = new gc_source{std::move(s)};
    return GC_OK;
}

int gc_source_create_synthetic_gpu(const gc_volume_cfg *vc, const gc_synth_params *p, int device, gc_source **out)
{
    if (!vc || !p || !out) return fail(GC_ERR_ARG, "null argument");
    auto s = std::make_unique<DeviceSynthSource>();
    s->init(*vc);
    s->synth = std::make_unique<Synth>(*p, vc->initial_sizes[0], vc->encoded_bytes);
    s->nthreads = p->threads ? p->threads : hw_threads();
    s->device = device;
    GC_TRY(s->setup());
    *out = new gc_source{std::move(s)};
    return GC_OK;
}

int gc_source_create_gpu_pyramid(const gc_volume_cfg *vc, const gc_synth_params *p, int device, uint32_t slab_planes,
                                 gc_source **out, double *build_ms)
{
    if (!vc || !p || !out) return fail(GC_ERR_ARG, "null argument");
    auto s = std::make_unique<DevicePyramidSource>();
    s->init(*vc);
    s->synth = std::make_unique<Synth>(*p, vc->initial_sizes[0], vc->encoded_bytes);
    s->nthreads = p->threads ? p->threads : hw_threads();
    s->device = device;
    GC_TRY(s->setup());
    GC_TRY(s->build(slab_planes ? slab_planes : 64u));
    if (build_ms) *build_ms = s->build_ms;
    *out = new gc_source{std::move(s)};
    return GC_OK;
}

int gc_source_get_device(gc_source *src, const uint32_t *keys, uint32_t n, void *d_dst, float threshold,
                         uint8_t *empty)
{
    if (!src || (n && (!keys || !d_dst || !empty))) return fail(GC_ERR_ARG, "null argument");
    Source &S = *src->impl;
    if (!S.on_device()) return fail(GC_ERR_ARG, "source does not write bricks on the device");
    hipStream_t s = nullptr;
    GC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const int rc = S.get_device(keys, n, static_cast<uint8_t *>(d_dst), threshold, empty, s);
    hipStreamDestroy(s);
    return rc;
}

int gc_source_create_callback(const gc_volume_cfg *vc, gc_brick_fn fn, void *user, gc_source **out)
{
    if (!vc || !fn || !out) return fail(GC_ERR_ARG, "null argument");
    auto s = std::make_unique<CallbackSource>();
    s->init(*vc);
    s->fn = fn;
    s->user = user;
    *out = new gc_source{std::move(s)};
    return GC_OK;
}

int gc_source_create_gpu_mip(const gc_volume_cfg *vc, gc_source *level0, int device, uint64_t memo_bytes,
                             gc_source **out)
{
    if (!vc || !level0 || !out) return fail(GC_ERR_ARG, "null argument");
    if (vc->nb_levels == 0 || vc->nb_levels > GC_MAX_LEVELS || vc->brick_size[0] <= 2 * vc->covering[0])
        return fail(GC_ERR_CONFIG, "bad volume configuration");
    if (vc->big_brick_size[0] != 1 || vc->big_brick_size[1] != 1 || vc->big_brick_size[2] != 1)
        return fail(GC_ERR_CONFIG, "mip on miss needs BigBrickSize 1");
    const Source &B = *level0->impl;
    if (B.bs != vc->brick_size[0] || B.cov != vc->covering[0] || B.bytes != vc->encoded_bytes ||
        B.BB[0] != 1 || B.BB[1] != 1 || B.BB[2] != 1)
        return fail(GC_ERR_CONFIG, "level-0 source has another brick layout");
    for (uint32_t l = 1; l < vc->nb_levels; ++l)
        for (int a = 0; a < 3; ++a)
            if (vc->initial_sizes[l][a] != vc->initial_sizes[l - 1][a] / 2)
                return fail(GC_ERR_CONFIG, "level sizes do not follow the 2x2x2 reduction");
    auto m = std::make_unique<MipSource>();
    m->init(*vc);
    m->base = level0->impl.get();
    m->device = device;
    m->memo_cap = memo_bytes;
    *out = new gc_source{std::move(m)};
    return GC_OK;
}

int gc_source_destroy(gc_source *src)
{
    delete src;
    return GC_OK;
}

int gc_source_get(gc_source *src, uint32_t level, const uint32_t big_brick[3], void *dst, int *status)
{
    if (!src || !big_brick || !dst) return fail(GC_ERR_ARG, "null argument");
    const int s = src->impl->get(level, big_brick, dst);
    if (status) *status = s;
    return GC_OK;
}

int gc_synth_volume(const gc_synth_params *p, const uint32_t dims[3], uint32_t encoded_bytes, void *host_out)
{
    if (!p || !dims || !host_out) return fail(GC_ERR_ARG, "null argument");
    Synth S(*p, dims, encoded_bytes);
    const uint32_t slab = 8;
    const uint32_t nslab = (dims[2] + slab - 1) / slab;
    const uint32_t nth = p->threads ? p->threads : hw_threads();
    std::memset(host_out, 0, (size_t)dims[0] * dims[1] * dims[2] * encoded_bytes);
    parallel_for(nslab, nth, [&](uint32_t i) {
        const int64_t o[3] = {0, 0, (int64_t)i * slab};
        const uint32_t n[3] = {dims[0], dims[1], std::min(slab, dims[2] - i * slab)};
        uint8_t *dst = static_cast<uint8_t *>(host_out) + (uint64_t)o[2] * dims[0] * dims[1] * encoded_bytes;
        S.fill_box(0, o, n, dims, dst);
    });
    return GC_OK;
}

int gc_cache_create(const gc_volume_cfg *vc, const gc_cache_cfg *cc, gc_source *src, int device, gc_cache **out)
{
    if (!vc || !cc || !src || !out) return fail(GC_ERR_ARG, "null argument");
    std::string why;
    if (!validate(*vc, *cc, why)) return fail(GC_ERR_CONFIG, why);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GC_ERR_NODEV, "no HIP device");
    if (device < 0 || device >= ndev) return fail(GC_ERR_ARG, "bad device id");
    auto c = std::make_unique<gc_cache>();
    c->impl.vc = *vc;
    c->impl.cc = *cc;
    c->impl.src = src->impl.get();
    c->impl.device = device;
    const int rc = c->impl.create();
    if (rc != GC_OK) return rc;
    *out = c.release();
    return GC_OK;
}

int gc_cache_destroy(gc_cache *c)
{
    delete c;
    return GC_OK;
}

int gc_cache_update(gc_cache *c, int *done)
{
    if (!c) return fail(GC_ERR_ARG, "null cache");
    GC_HIP(hipSetDevice(c->impl.device));
    return c->impl.update(done);
}

int gc_cache_completude(gc_cache *c, float *out)
{
    if (!c || !out) return fail(GC_ERR_ARG, "null argument");
    Cache &C = c->impl;
    GC_HIP(launch_occupancy(C.meta, C.g.nslots, C.d_count, C.stream));
    uint32_t cnt = 0;
    GC_HIP(hipMemcpyAsync(&cnt, C.d_count, 4, hipMemcpyDeviceToHost, C.stream));
    GC_HIP(hipStreamSynchronize(C.stream));
    *out = (float)cnt / (float)C.g.nslots;
    return GC_OK;
}

int gc_cache_load_all(gc_cache *c, const void *const *d_levels)
{
    if (!c || !d_levels) return fail(GC_ERR_ARG, "null argument");
    Cache &C = c->impl;
    if (C.cc.empty_threshold > 0.f)
        return fail(GC_ERR_CONFIG, "load_all maps every brick and does not classify empty bricks");
    if (C.vc.big_brick_size[0] != 1 || C.vc.big_brick_size[1] != 1 || C.vc.big_brick_size[2] != 1)
        return fail(GC_ERR_CONFIG, "load_all needs BigBrickSize 1");
    std::vector<uint32_t> asked;
    const gc_geometry &g = C.g;
    for (uint32_t z = 0; z < g.mrpd_dim[2]; ++z)
        for (uint32_t y = 0; y < g.mrpd_dim[1]; ++y)
            for (uint32_t x = 0; x < g.mrpd_dim[0]; ++x) {
                uint32_t level = 0;
                while (x >= g.level_coord[level] + g.level_dim[level][0]) ++level;
                if (y >= g.level_dim[level][1] || z >= g.level_dim[level][2]) continue;
                asked.insert(asked.end(), {x - g.level_coord[level], y, z, level});
            }
    const uint32_t n = (uint32_t)(asked.size() / 4);
    if (n > g.nslots) return fail(GC_ERR_CONFIG, "pool too small to hold every brick");
    uint32_t *d_asked = nullptr, *d_dims = nullptr;
    const void **d_lv = nullptr;
    GC_TRY(dalloc(&d_asked, asked.size() * 4));
    GC_TRY(dalloc(&d_dims, (size_t)g.nb_levels * 12));
    GC_TRY(dalloc(&d_lv, (size_t)g.nb_levels * sizeof(void *)));
    GC_HIP(hipMemcpyAsync(d_asked, asked.data(), asked.size() * 4, hipMemcpyHostToDevice, C.stream));
    GC_HIP(hipMemcpyAsync(d_dims, C.vc.initial_sizes, (size_t)g.nb_levels * 12, hipMemcpyHostToDevice, C.stream));
    GC_HIP(hipMemcpyAsync(d_lv, d_levels, (size_t)g.nb_levels * sizeof(void *), hipMemcpyHostToDevice, C.stream));
    if (g.nb_tables) {
        GC_HIP(launch_deref_tables(C.cells, C.tabs, C.order, g.nslots, n, C.meta, C.d_geom, C.ts, C.stream));
        GC_HIP(launch_assign_cut(C.order, g.nslots, n, d_asked, C.meta, C.pool, d_lv, d_dims, C.d_geom,
                                 C.vc.encoded_bytes, C.stream));
        GC_HIP(launch_reference_tables(C.mrpd, C.cells, C.tabs, C.order, g.nslots, n, n, d_asked, C.meta, C.d_geom,
                                       C.ts, C.d_unref, C.stream));
    } else {
        GC_HIP(launch_deref(C.mrpd, C.cells, C.order, g.nslots, n, C.meta, C.d_geom, C.stream));
        GC_HIP(launch_assign_cut(C.order, g.nslots, n, d_asked, C.meta, C.pool, d_lv, d_dims, C.d_geom,
                                 C.vc.encoded_bytes, C.stream));
        int serial = 0;
        GC_TRY(C.insert_serial_needed(asked.data(), n, serial));
        GC_HIP(launch_insert(C.mrpd, C.cells, C.order, g.nslots, n, n, d_asked, C.d_geom, serial, C.stream));
    }
    GC_HIP(hipStreamSynchronize(C.stream));
    hipFree(d_asked);
    hipFree(d_dims);
    hipFree(d_lv);
    return GC_OK;
}

int gc_cache_timestamp(gc_cache *c, uint32_t *ts)
{
    if (!c || !ts) return fail(GC_ERR_ARG, "null argument");
    *ts = c->impl.ts;
    return GC_OK;
}

int gc_cache_counters_get(gc_cache *c, gc_cache_counters *out)
{
    if (!c || !out) return fail(GC_ERR_ARG, "null argument");
    *out = c->impl.counters;
    out->slots = c->impl.g.nslots;
    out->fetch_ms = (double)c->impl.fetch_ns.load() / 1e6;
    return GC_OK;
}

int gc_cache_stream(gc_cache *c, void **stream)
{
    if (!c || !stream) return fail(GC_ERR_ARG, "null argument");
    *stream = c->impl.stream;
    return GC_OK;
}

int gc_cache_view_get(gc_cache *c, gc_cache_view *out)
{
    if (!c || !out) return fail(GC_ERR_ARG, "null argument");
    Cache &C = c->impl;
    out->mrpd = C.mrpd;
    out->request = C.request;
    out->usage = C.usage;
    out->pool = C.pool;
    out->geometry = C.d_geom;
    out->timestamp = C.ts;
    out->brick_size = C.g.brick_size;
    out->covering = C.g.covering;
    out->encoded_bytes = C.g.encoded_bytes;
    out->nb_tables = C.g.nb_tables;
    for (int i = 0; i < GC_MAX_TABLES; ++i) {
        out->table_words[i] = C.tabs.words[i];
        out->table_usage[i] = C.tabs.usage[i];
    }
    return GC_OK;
}

int gc_cache_geometry(gc_cache *c, gc_geometry *out)
{
    if (!c || !out) return fail(GC_ERR_ARG, "null argument");
    *out = c->impl.g;
    return GC_OK;
}

int gc_geometry_compute(const gc_volume_cfg *vc, const gc_cache_cfg *cc, gc_geometry *out)
{
    if (!vc || !cc || !out) return fail(GC_ERR_ARG, "null argument");
    std::string why;
    if (!validate(*vc, *cc, why)) return fail(GC_ERR_CONFIG, why);
    compute_geometry(*vc, *cc, *out);
    return GC_OK;
}

static int d2h_sync(Cache &C, void *dst, const void *src, size_t bytes)
{
    GC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, C.stream));
    GC_HIP(hipStreamSynchronize(C.stream));
    return GC_OK;
}
static int h2d_sync(Cache &C, void *dst, const void *src, size_t bytes)
{
    GC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, C.stream));
    GC_HIP(hipStreamSynchronize(C.stream));
    return GC_OK;
}

int gc_cache_read_mrpd(gc_cache *c, uint32_t *host)
{
    if (!c || !host) return fail(GC_ERR_ARG, "null argument");
    return d2h_sync(c->impl, host, c->impl.mrpd, c->impl.g.top_cells * 16);
}

int gc_cache_read_cells(gc_cache *c, uint32_t *host)
{
    if (!c || !host) return fail(GC_ERR_ARG, "null argument");
    return d2h_sync(c->impl, host, c->impl.cells, c->impl.g.top_cells * 4);
}
int gc_cache_read_request(gc_cache *c, uint32_t *host)
{
    if (!c || !host) return fail(GC_ERR_ARG, "null argument");
    return d2h_sync(c->impl, host, c->impl.request, c->impl.g.mrpd_cells * 4);
}
int gc_cache_write_request(gc_cache *c, const uint32_t *host)
{
    if (!c || !host) return fail(GC_ERR_ARG, "null argument");
    return h2d_sync(c->impl, c->impl.request, host, c->impl.g.mrpd_cells * 4);
}
int gc_cache_read_usage(gc_cache *c, uint32_t *host)
{
    if (!c || !host) return fail(GC_ERR_ARG, "null argument");
    return d2h_sync(c->impl, host, c->impl.usage, (size_t)c->impl.g.nslots * 4);
}
int gc_cache_write_usage(gc_cache *c, const uint32_t *host)
{
    if (!c || !host) return fail(GC_ERR_ARG, "null argument");
    return h2d_sync(c->impl, c->impl.usage, host, (size_t)c->impl.g.nslots * 4);
}

int gc_cache_read_lru(gc_cache *c, gc_lru_entry *host)
{
    if (!c || !host) return fail(GC_ERR_ARG, "null argument");
    Cache &C = c->impl;
    std::vector<uint32_t> ord(C.g.nslots);
    std::vector<SlotMeta> mt(C.g.nslots);
    GC_TRY(d2h_sync(C, ord.data(), C.order, ord.size() * 4));
    GC_TRY(d2h_sync(C, mt.data(), C.meta, mt.size() * sizeof(SlotMeta)));
    for (uint32_t i = 0; i < C.g.nslots; ++i) {
        slot_origin(C.g, ord[i], host[i].cache_pos);
        host[i].level = mt[ord[i]].level;
        std::memcpy(host[i].vol_pos, mt[ord[i]].vol, sizeof(host[i].vol_pos));
    }
    return GC_OK;
}

static int table_arg(gc_cache *c, uint32_t i, const void *host)
{
    if (!c || !host) return fail(GC_ERR_ARG, "null argument");
    if (i >= c->impl.g.nb_tables) return fail(GC_ERR_ARG, "no such page-table cache");
    return GC_OK;
}

int gc_cache_read_table(gc_cache *c, uint32_t i, uint32_t *host4)
{
    GC_TRY(table_arg(c, i, host4));
    Cache &C = c->impl;
    const gc_table_geometry &t = C.g.table[i];
    const uint32_t b3 = table_b3(t);
    std::vector<uint32_t> w((size_t)t.nblocks * b3);
    GC_TRY(d2h_sync(C, w.data(), C.tabs.words[i], w.size() * 4));
    // the reference's surface: entry (x, y, z) of block (x, y, z) / block;
    // a word names a block of the next table or a brick slot
    const bool last = i + 1 == C.g.nb_tables;
    for (uint32_t b = 0; b < t.nblocks; ++b) {
        uint32_t o[3];
        block_origin(t, b, o);
        for (uint32_t k = 0; k < b3; ++k) {
            const uint32_t lx = k % t.block[0], ly = (k / t.block[0]) % t.block[1], lz = k / (t.block[0] * t.block[1]);
            const uint64_t e = (o[0] + lx) + (uint64_t)t.surf[0] * ((o[1] + ly) + (uint64_t)t.surf[1] * (o[2] + lz));
            const uint32_t word = w[(size_t)b * b3 + k], st = word & 3u;
            uint32_t co[3] = {0, 0, 0};
            if (st == 1) {
                if (last) slot_origin(C.g, word >> 2, co);
                else block_origin(C.g.table[i + 1], word >> 2, co);
            }
            host4[4 * e] = co[0];
            host4[4 * e + 1] = co[1];
            host4[4 * e + 2] = co[2];
            host4[4 * e + 3] = st;
        }
    }
    return GC_OK;
}

int gc_cache_read_table_usage(gc_cache *c, uint32_t i, uint32_t *host)
{
    GC_TRY(table_arg(c, i, host));
    return d2h_sync(c->impl, host, c->impl.tabs.usage[i], (size_t)c->impl.g.table[i].nblocks * 4);
}

int gc_cache_write_table_usage(gc_cache *c, uint32_t i, const uint32_t *host)
{
    GC_TRY(table_arg(c, i, host));
    return h2d_sync(c->impl, c->impl.tabs.usage[i], host, (size_t)c->impl.g.table[i].nblocks * 4);
}

int gc_cache_read_table_lru(gc_cache *c, uint32_t i, gc_lru_entry *host)
{
    GC_TRY(table_arg(c, i, host));
    Cache &C = c->impl;
    const gc_table_geometry &t = C.g.table[i];
    std::vector<uint32_t> ord(t.nblocks);
    std::vector<SlotMeta> mt(t.nblocks);
    GC_TRY(d2h_sync(C, ord.data(), C.tabs.order[i], ord.size() * 4));
    GC_TRY(d2h_sync(C, mt.data(), C.tabs.meta[i], mt.size() * sizeof(SlotMeta)));
    for (uint32_t k = 0; k < t.nblocks; ++k) {
        block_origin(t, ord[k], host[k].cache_pos);
        host[k].level = mt[ord[k]].level;
        std::memcpy(host[k].vol_pos, mt[ord[k]].vol, sizeof(host[k].vol_pos));
    }
    return GC_OK;
}

int gc_cache_unreferenced(gc_cache *c, uint32_t *out)
{
    if (!c || !out) return fail(GC_ERR_ARG, "null argument");
    return d2h_sync(c->impl, out, c->impl.d_unref, 4);
}

int gc_cache_read_pool(gc_cache *c, void *host)
{
    if (!c || !host) return fail(GC_ERR_ARG, "null argument");
    Cache &C = c->impl;
    const uint32_t bs = C.g.brick_size, B = C.vc.encoded_bytes;
    const uint64_t bvox = (uint64_t)bs * bs * bs;
    std::vector<uint8_t> tmp((size_t)C.g.nslots * bvox * B);
    GC_TRY(d2h_sync(C, tmp.data(), C.pool, tmp.size()));
    const uint64_t PX = (uint64_t)C.g.cache_size[0] * bs, PY = (uint64_t)C.g.cache_size[1] * bs;
    for (uint32_t s = 0; s < C.g.nslots; ++s) {
        uint32_t o[3];
        slot_origin(C.g, s, o);
        for (uint32_t z = 0; z < bs; ++z)
            for (uint32_t y = 0; y < bs; ++y) {
                const uint64_t d = (o[0] + PX * ((uint64_t)o[1] + y + PY * ((uint64_t)o[2] + z))) * B;
                const uint64_t so = ((uint64_t)s * bvox + (uint64_t)bs * (y + (uint64_t)bs * z)) * B;
                std::memcpy(static_cast<uint8_t *>(host) + d, tmp.data() + so, (size_t)bs * B);
            }
    }
    return GC_OK;
}

int gc_cache_read_asked(gc_cache *c, uint32_t *host4, uint32_t *nb_asked, uint32_t *nb_non_empty)
{
    if (!c) return fail(GC_ERR_ARG, "null cache");
    Cache &C = c->impl;
    if (C.handler.upload_pending) GC_HIP(hipEventSynchronize(C.uploaded));
    if (host4) std::memcpy(host4, C.h_meta, (size_t)C.handler.n * 16);
    if (nb_asked) *nb_asked = C.handler.n;
    if (nb_non_empty) *nb_non_empty = C.handler.nne;
    return GC_OK;
}

int gc_frame_create(uint32_t width, uint32_t height, uint32_t precision, gc_frame **out)
{
    if (!out || width == 0 || height == 0) return fail(GC_ERR_ARG, "bad frame size");
    if (precision != 32 && precision != 64) return fail(GC_ERR_ARG, "precision must be 32 or 64");
    auto f = std::make_unique<gc_frame>();
    Frame &F = f->impl;
    F.w = width; F.h = height; F.precision = precision;
    if ((uint64_t)width * height >= (1ull << DVR_PIXEL_BITS)) return fail(GC_ERR_ARG, "frame larger than 2^26 pixels");
    GC_TRY(dalloc(&F.pixels, (size_t)width * height * 4));
    GC_TRY(dalloc(&F.tbuf, (size_t)width * height * 4));
    GC_TRY(dalloc(&F.cstate, F.cbytes()));
    GC_TRY(dalloc(&F.rpos, (size_t)width * height * sizeof(float4)));
    GC_TRY(dalloc(&F.qbuf, (size_t)width * height * sizeof(double)));
    GC_TRY(dalloc(&F.ctl, 64));
    GC_TRY(dalloc(&F.lists[0], (size_t)width * height * 4));
    GC_TRY(dalloc(&F.lists[1], (size_t)width * height * 4));
    GC_TRY(dalloc(&F.handed, (size_t)width * height * 4));
    GC_TRY(dalloc(&F.handed_pos, (size_t)width * height * sizeof(float4)));
    GC_TRY(dalloc(&F.hand_ctl, HAND_CTL_WORDS * 4));
    GC_TRY(dalloc(&F.stop, ((size_t)width * height + 3) & ~(size_t)3));
    GC_HIP(hipMemset(F.stop, 0, ((size_t)width * height + 3) & ~(size_t)3));
    GC_HIP(hipMemset(F.ctl, 0, 64));
    GC_HIP(hipMemset(F.pixels, 0, (size_t)width * height * 4));
    GC_HIP(hipMemset(F.tbuf, 0, (size_t)width * height * 4));
    GC_HIP(hipMemset(F.cstate, 0, F.cbytes()));
    *out = f.release();
    return GC_OK;
}

int gc_frame_destroy(gc_frame *f)
{
    delete f;
    return GC_OK;
}

int gc_frame_reset(gc_frame *f, void *stream)
{
    if (!f) return fail(GC_ERR_ARG, "null frame");
    Frame &F = f->impl;
    if (!stream) {
        // no stream given: reset synchronously so that any stream sees it
        GC_HIP(hipMemset(F.tbuf, 0, (size_t)F.w * F.h * 4));
        GC_HIP(hipMemset(F.cstate, 0, F.cbytes()));
        GC_HIP(hipDeviceSynchronize());
        return GC_OK;
    }
    // fill kernels rather than memsets: a caller may capture this stream
    GC_HIP(launch_fill32(reinterpret_cast<uint32_t *>(F.tbuf), F.w * F.h, 0, (hipStream_t)stream));
    GC_HIP(launch_fill32(static_cast<uint32_t *>(F.cstate), (uint32_t)(F.cbytes() / 4), 0, (hipStream_t)stream));
    return GC_OK;
}

int gc_frame_read_pixels(gc_frame *f, uint32_t *host)
{
    if (!f || !host) return fail(GC_ERR_ARG, "null argument");
    GC_HIP(hipDeviceSynchronize());
    GC_HIP(hipMemcpy(host, f->impl.pixels, (size_t)f->impl.w * f->impl.h * 4, hipMemcpyDeviceToHost));
    return GC_OK;
}

int gc_frame_read_state(gc_frame *f, float *tbuf, void *cstate)
{
    if (!f) return fail(GC_ERR_ARG, "null frame");
    GC_HIP(hipDeviceSynchronize());
    if (tbuf) GC_HIP(hipMemcpy(tbuf, f->impl.tbuf, (size_t)f->impl.w * f->impl.h * 4, hipMemcpyDeviceToHost));
    if (cstate) GC_HIP(hipMemcpy(cstate, f->impl.cstate, f->impl.cbytes(), hipMemcpyDeviceToHost));
    return GC_OK;
}

int gc_frame_buffers(gc_frame *f, uint32_t **pixels, float **tbuf, void **cstate)
{
    if (!f) return fail(GC_ERR_ARG, "null frame");
    if (pixels) *pixels = f->impl.pixels;
    if (tbuf) *tbuf = f->impl.tbuf;
    if (cstate) *cstate = f->impl.cstate;
    return GC_OK;
}

int gc_render_dvr(gc_cache *c, const gc_view *view, const float *d_tf, uint32_t tf_size, gc_frame *f,
                  gc_render_stats *d_stats, void *stream)
{
    if (!c || !view || !d_tf || !f || tf_size == 0) return fail(GC_ERR_ARG, "bad argument");
    Cache &C = c->impl;
    Frame &F = f->impl;
    if (view->width != F.w || view->height != F.h) return fail(GC_ERR_ARG, "view and frame sizes differ");
    if (view->precision != F.precision) return fail(GC_ERR_ARG, "view and frame precision differ");
    if (view->tile >= GC_TILES) return fail(GC_ERR_ARG, "tile must be 0..5");
    if (view->compositing > 1) return fail(GC_ERR_ARG, "compositing must be 0 (DVR) or 1 (MOP)");
    DvrLaunch L{};
    launch_base(C, L);
    view_of(*view, d_tf, tf_size, F, L.kexp[0], L.view0);
    if (view->compositing == 1) {
        // maximum projection: one launch over the 8x8 tiles
        L.n_views = 1;
        L.n_packets = tile_packets(0, view->width, view->height);
        L.stats = d_stats;
        hipStream_t s = stream ? (hipStream_t)stream : C.stream;
        GC_HIP(launch_mop(L, C.vc.encoded_bytes, view->precision, 0, s));
        return GC_OK;
    }
    L.exact_resume = L.view0.rpos ? 1u : 0u;
    finish_exact(L, &L.view0, 1);
    L.tf_lds_entries = dvr_assign_tf_lds(&L.view0, 1, C.vc.encoded_bytes, &L.tf_lds_tables);
    L.n_views = 1;
    L.n_packets = packets_of(*view);
    L.stats = d_stats;
    L.queue = F.ctl;
    L.refill_min = view->refill_min ? view->refill_min : 16;
    L.march = view->march ? view->march : 8;
    if (L.march > 8) return fail(GC_ERR_ARG, "march must be 1..8");
    hipStream_t s = stream ? (hipStream_t)stream : C.stream;
    GC_HIP(launch_fill32(F.ctl, 8, 0, s));
    const uint32_t grid = dvr_grid(C.vc.encoded_bytes, view->precision, view->tile, L.n_packets,
                                   dvr_lds_bytes(L.nb_lods, 1, L.tf_lds_entries), L.tf_lds_tables);
    GC_HIP(launch_dvr(L, C.vc.encoded_bytes, view->precision, view->tile, grid, s));
    return GC_OK;
}

// One launch of the progressive frame loop: the first covers every tile of
// the view, the next ones only the rays the previous launch stopped at a miss
// (the other rays are finished and would neither sample nor write anything
// new).  Returns the index of the list the launch fills.
// One launch of a frame.  d_stats[0] gathers the counters of tile / lane-
// per-ray launches (k_dvr), d_stats[1] those of resumed-ray launches
// (k_dvr_ray); *lane tells which one ran.
static int frame_launch(Cache &C, Frame &F, const gc_view &view, const float *d_tf, uint32_t tf_size,
                        gc_render_stats *d_stats, int first, int &cur, hipStream_t s, uint32_t list_rays, int *lane,
                        bool *handed)
{
    *handed = false;
    DvrLaunch L{};
    launch_base(C, L);
    view_of(view, d_tf, tf_size, F, L.kexp[0], L.view0);
    L.view0.stop = F.stop;
    L.view0.stop_tile_off = 0;
    L.exact_resume = L.view0.rpos ? 1u : 0u;
    finish_exact(L, &L.view0, 1);
    L.tf_lds_entries = dvr_assign_tf_lds(&L.view0, 1, C.vc.encoded_bytes, &L.tf_lds_tables);
    L.n_views = 1;
    L.n_packets = first ? packets_of(view) : 0;
    L.stats = d_stats;
    L.queue = F.ctl;
    L.refill_min = view.refill_min ? view.refill_min : 16;
    L.march = view.march ? view.march : 8;
    if (L.march > 8) return fail(GC_ERR_ARG, "march must be 1..8");
    const int out = first ? 0 : 1 - cur;
    if (!first) {
        L.list = F.lists[cur];
        L.list_count = F.ctl + 8 + cur;
    }
    GC_HIP(launch_fill32(F.ctl, 8, 0, s));
    GC_HIP(launch_fill32(F.ctl + 8 + out, 1, 0, s));
    *lane = view.compositing == 1 || !(!first && list_rays <= C.ray_list_max && dvr_ray_supported(L, C.vc.encoded_bytes));
    if (view.compositing == 1) {
        if (first) L.n_packets = tile_packets(0, view.width, view.height);
        GC_HIP(launch_mop(L, C.vc.encoded_bytes, view.precision, first ? 0u : list_rays, s));
    } else if (!*lane) {
        // a short list of resumed rays: a group of lanes per ray
        L.stats = d_stats + 1;
        const uint32_t group = C.group_for(list_rays, L.pt.n != 0);
        GC_HIP(launch_dvr_ray(L, C.vc.encoded_bytes, view.precision, list_rays,
                              dvr_ray_grid(C.vc.encoded_bytes, view.precision, list_rays,
                                           dvr_ray_lds_bytes(1, L.tf_lds_entries), group),
                              s, group));
    } else {
        const uint32_t grid = dvr_grid(C.vc.encoded_bytes, view.precision, view.tile,
                                       first ? L.n_packets : (F.w * F.h + 63) / 64,
                                       dvr_lds_bytes(L.nb_lods, 1, L.tf_lds_entries), L.tf_lds_tables);
        // The launch lasts as long as its longest rays, which the lane march
        // walks one sample per step: once most of its waves have run out of
        // rays, the rays left are handed over to a group march that continues
        // them (same state, same stamps and counters) before the cache update.
        const bool handoff = handoff_possible(C, L);
        DvrLaunch H{};
        if (handoff) {
            handoff_launches(C, L, H, grid, F.handed, F.handed_pos, F.hand_ctl);
            GC_HIP(launch_fill32(F.hand_ctl, HAND_CTL_WORDS, 0, s));
            *handed = true;
        }
        GC_HIP(launch_dvr(L, C.vc.encoded_bytes, view.precision, view.tile, grid, s));
        if (handoff) {
            const uint32_t group = handoff_group_of(C, L);
            GC_HIP(launch_dvr_ray(H, C.vc.encoded_bytes, view.precision, F.w * F.h,
                                  dvr_ray_grid(C.vc.encoded_bytes, view.precision, F.w * F.h,
                                               dvr_ray_lds_bytes(1, H.tf_lds_entries), group),
                                  s, group));
        }
    }
    // the stopped rays, in screen order, form the next launch's list
    CollectArgs ca{};
    ca.one = L.view0;
    ca.n_views = 1;
    ca.n_tiles = stop_tiles(F.w, F.h);
    ca.out_list = F.lists[out];
    ca.out_count = F.ctl + 8 + out;
    GC_HIP(launch_collect_stopped(ca, s));
    cur = out;
    return GC_OK;
}

int gc_cache_busy(gc_cache *c, int *busy)
{
    if (!c || !busy) return fail(GC_ERR_ARG, "null argument");
    *busy = c->impl.handler.new_request.load() ? 1 : 0;
    return GC_OK;
}

int gc_cache_wait_idle(gc_cache *c)
{
    if (!c) return fail(GC_ERR_ARG, "null cache");
    Handler &H = c->impl.handler;
    std::unique_lock<std::mutex> lk(H.m);
    H.idle_cv.wait(lk, [&] { return !H.new_request.load(); });
    return GC_OK;
}

int gc_render_frame(gc_cache *c, const gc_view *view, const float *d_tf, uint32_t tf_size, gc_frame *f,
                    uint32_t max_launches, gc_frame_report *report)
{
    if (!c || !view || !d_tf || !f || !report) return fail(GC_ERR_ARG, "null argument");
    Cache &C = c->impl;
    GC_HIP(hipSetDevice(C.device));
    std::memset(report, 0, sizeof(*report));
    if (!C.d_frame_stats) {
        GC_TRY(dalloc(&C.d_frame_stats, 2 * sizeof(gc_render_stats)));
        GC_HIP(hipEventCreate(&C.ev0));
        GC_HIP(hipEventCreate(&C.ev1));
    }
    if (view->width != f->impl.w || view->height != f->impl.h || view->precision != f->impl.precision ||
        view->tile >= GC_TILES || view->compositing > 1)
        return fail(GC_ERR_ARG, "view does not match the frame");
    GC_HIP(hipMemsetAsync(C.d_frame_stats, 0, 2 * sizeof(gc_render_stats), C.stream));
    if (C.log) *C.log_count = 0;
    gc_render_stats logged[2] = {};   // the frame's counters at the last logged launch
    int cur = 0;
    for (uint32_t it = 0; it < max_launches; ++it) {
        uint32_t list_rays = 0;
        if (it > 0)
            GC_HIP(hipMemcpyAsync(&list_rays, f->impl.ctl + 8 + cur, 4, hipMemcpyDeviceToHost, C.stream));
        GC_HIP(hipStreamSynchronize(C.stream));
        GC_HIP(hipEventRecord(C.ev0, C.stream));
        int lane = 1;
        bool handed = false;
        GC_TRY(frame_launch(C, f->impl, *view, d_tf, tf_size, C.d_frame_stats, it == 0, cur, C.stream, list_rays,
                            &lane, &handed));
        GC_HIP(hipEventRecord(C.ev1, C.stream));
        ++report->launches;
        report->lane_launches += (uint32_t)lane;
        int done = 0;
        const auto t0 = std::chrono::steady_clock::now();
        GC_TRY(C.update(&done));
        ++report->updates;
        GC_HIP(hipEventSynchronize(C.ev1));
        float ms = 0.f;
        GC_HIP(hipEventElapsedTime(&ms, C.ev0, C.ev1));
        report->kernel_ms += ms;
        if (lane) report->lane_kernel_ms += ms;
        gc_launch_record *rec = nullptr;
        if (C.log && *C.log_count < C.log_cap) {
            gc_render_stats st[2];
            uint32_t n_handed = 0;
            GC_HIP(hipMemcpyAsync(st, C.d_frame_stats, sizeof(st), hipMemcpyDeviceToHost, C.stream));
            if (handed) GC_HIP(hipMemcpyAsync(&n_handed, f->impl.hand_ctl, 4, hipMemcpyDeviceToHost, C.stream));
            GC_HIP(hipStreamSynchronize(C.stream));
            rec = &C.log[(*C.log_count)++];
            std::memset(rec, 0, sizeof(*rec));
            rec->launch = it;
            rec->kernel = view->compositing == 1 ? 2u : (lane ? 0u : 1u);
            rec->list_rays = list_rays;
            rec->group = view->compositing != 1 && !lane ? C.group_for(list_rays, C.g.nb_tables != 0) : 0u;
            const int k = lane ? 0 : 1;
            rec->stats.mapped_samples = st[k].mapped_samples - logged[k].mapped_samples;
            rec->stats.brick_visits = st[k].brick_visits - logged[k].brick_visits;
            rec->stats.pixels = st[k].pixels - logged[k].pixels;
            rec->stats.unmapped_rays = st[k].unmapped_rays - logged[k].unmapped_rays;
            rec->kernel_ms = ms;
            rec->handed_over = n_handed;
            logged[0] = st[0];
            logged[1] = st[1];
        }
        if (done) {
            report->update_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            report->complete = 1;
            break;
        }
        // a request was issued: wait for it and load it before the next launch
        GC_TRY(gc_cache_wait_idle(c));
        const uint32_t before = C.handler.nne;
        GC_TRY(C.update(&done));
        ++report->updates;
        report->bricks_loaded += before;
        if (rec) rec->bricks_loaded = before;
        report->update_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    gc_render_stats st[2];
    GC_HIP(hipMemcpyAsync(st, C.d_frame_stats, sizeof(st), hipMemcpyDeviceToHost, C.stream));
    GC_HIP(hipStreamSynchronize(C.stream));
    report->lane_stats = st[0];
    report->stats.mapped_samples = st[0].mapped_samples + st[1].mapped_samples;
    report->stats.brick_visits = st[0].brick_visits + st[1].brick_visits;
    report->stats.pixels = st[0].pixels + st[1].pixels;
    report->stats.unmapped_rays = st[0].unmapped_rays + st[1].unmapped_rays;
    return GC_OK;
}

int gc_cache_set_ray_list_max(gc_cache *c, uint32_t max_rays)
{
    if (!c) return fail(GC_ERR_ARG, "null cache");
    c->impl.ray_list_max = max_rays;
    return GC_OK;
}

int gc_cache_set_launch_log(gc_cache *c, gc_launch_record *records, uint32_t capacity, uint32_t *count)
{
    if (!c) return fail(GC_ERR_ARG, "null cache");
    if (records && !count) return fail(GC_ERR_ARG, "launch log without a count");
    c->impl.log = records;
    c->impl.log_cap = records ? capacity : 0;
    c->impl.log_count = records ? count : nullptr;
    if (count) *count = 0;
    return GC_OK;
}

int gc_cache_set_ray_handoff(gc_cache *c, uint32_t samples, uint32_t lanes, uint32_t exited_pct)
{
    if (!c) return fail(GC_ERR_ARG, "null cache");
    if (lanes != 0 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 64)
        return fail(GC_ERR_ARG, "lanes per ray must be 0, 4, 8, 16 or 64");
    c->impl.handoff_after = samples;
    if (exited_pct > 100) return fail(GC_ERR_ARG, "exited_pct must be 0..100");
    c->impl.handoff_group = lanes ? lanes : 16u;
    c->impl.handoff_exited_pct = exited_pct;
    return GC_OK;
}

int gc_cache_get_ray_handoff(const gc_cache *c, uint32_t *samples, uint32_t *lanes, uint32_t *exited_pct)
{
    if (!c || !samples || !lanes || !exited_pct) return fail(GC_ERR_ARG, "null argument");
    *samples = c->impl.handoff_after;
    *lanes = c->impl.handoff_group;
    *exited_pct = c->impl.handoff_exited_pct;
    return GC_OK;
}

int gc_cache_set_ray_group(gc_cache *c, uint32_t lanes)
{
    if (!c) return fail(GC_ERR_ARG, "null cache");
    if (lanes != 0 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 64)
        return fail(GC_ERR_ARG, "lanes per ray must be 0, 4, 8, 16 or 64");
    c->impl.ray_group = lanes;
    return GC_OK;
}

int gc_cache_set_max_requests(gc_cache *c, uint32_t max_requests)
{
    if (!c || max_requests == 0) return fail(GC_ERR_ARG, "bad argument");
    Cache &C = c->impl;
    if (max_requests > C.cap)
        return fail(GC_ERR_ARG, "max_requests above the capacity the cache was created with (" + std::to_string(C.cap) + ")");
    C.limit = max_requests;
    return GC_OK;
}

int gc_batch_create(gc_cache *c, uint32_t max_views, uint64_t max_pixels, uint32_t precision, gc_batch **out)
{
    if (!c || !out) return fail(GC_ERR_ARG, "null argument");
    if (max_views == 0 || max_views > DVR_MAX_VIEWS) return fail(GC_ERR_ARG, "max_views must be 1..64");
    if (precision != 32 && precision != 64) return fail(GC_ERR_ARG, "precision must be 32 or 64");
    if (max_pixels == 0 || max_pixels * (precision == 64 ? 8 : 4) >= (1ull << 32))
        return fail(GC_ERR_ARG, "max_pixels must be positive and below 2^29");
    auto b = std::make_unique<gc_batch>();
    Batch &B = b->impl;
    B.C = &c->impl;
    B.h = c;
    B.precision = precision;
    B.max_views = max_views;
    B.max_pixels = max_pixels;
    B.slots.resize(max_views);
    GC_HIP(hipSetDevice(B.C->device));
    GC_TRY(dalloc(&B.pixels, max_pixels * 4));
    GC_TRY(dalloc(&B.tbuf, max_pixels * 4));
    GC_TRY(dalloc(&B.cstate, max_pixels * B.cbytes_per_pixel()));
    GC_TRY(dalloc(&B.rpos, max_pixels * sizeof(float4)));
    GC_TRY(dalloc(&B.qbuf, max_pixels * sizeof(double)));
    GC_TRY(dalloc(&B.lists[0], max_pixels * 4));
    GC_TRY(dalloc(&B.lists[1], max_pixels * 4));
    GC_TRY(dalloc(&B.handed, max_pixels * 4));
    GC_TRY(dalloc(&B.handed_pos, max_pixels * sizeof(float4)));
    GC_TRY(dalloc(&B.hand_ctl, HAND_CTL_WORDS * 4));
    GC_TRY(dalloc(&B.stop, (max_pixels + 3) & ~3ull));
    GC_HIP(hipMemset(B.stop, 0, (max_pixels + 3) & ~3ull));
    GC_TRY(dalloc(&B.ctl, (16 + DVR_MAX_VIEWS) * 4));
    GC_TRY(dalloc(&B.d_views, (size_t)max_views * sizeof(DvrView)));
    GC_TRY(dalloc(&B.d_stats, sizeof(gc_render_stats)));
    GC_TRY(halloc(&B.h_views, (size_t)max_views * sizeof(DvrView)));
    GC_TRY(halloc(&B.h_stopped, (size_t)DVR_MAX_VIEWS * 4));
    std::memset(B.h_stopped, 0, (size_t)DVR_MAX_VIEWS * 4);
    GC_HIP(hipEventCreate(&B.ev0));
    GC_HIP(hipEventCreate(&B.ev1));
    GC_HIP(hipEventCreateWithFlags(&B.views_free, hipEventDisableTiming));
    GC_HIP(hipEventCreateWithFlags(&B.launched, hipEventDisableTiming));
    GC_HIP(hipEventRecord(B.views_free, B.C->stream));
    GC_HIP(hipEventRecord(B.launched, B.C->stream));
    GC_HIP(hipMemset(B.ctl, 0, (16 + DVR_MAX_VIEWS) * 4));
    GC_HIP(hipMemset(B.pixels, 0, max_pixels * 4));
    GC_HIP(hipMemset(B.d_stats, 0, sizeof(gc_render_stats)));
    *out = b.release();
    return GC_OK;
}

int gc_batch_destroy(gc_batch *b)
{
    if (!b) return GC_OK;
    hipSetDevice(b->impl.C->device);
    hipStreamSynchronize(b->impl.C->stream);
    delete b;
    return GC_OK;
}

int gc_batch_set_graph(gc_batch *b, int enable)
{
    if (!b) return fail(GC_ERR_ARG, "null batch");
    Batch &B = b->impl;
    B.use_graph = enable != 0;
    if (!B.use_graph) {
        GC_HIP(hipStreamSynchronize(B.C->stream));
        B.drop_graphs();
    }
    return GC_OK;
}

int gc_batch_set_launch(gc_batch *b, uint32_t tile, uint32_t march, uint32_t refill_min)
{
    if (!b) return fail(GC_ERR_ARG, "null batch");
    if (tile >= GC_TILES) return fail(GC_ERR_ARG, "tile must be 0..5");
    if (march > 8) return fail(GC_ERR_ARG, "march must be 1..8");
    if (refill_min > 64) return fail(GC_ERR_ARG, "refill_min must be 0..64");
    Batch &B = b->impl;
    B.tile = tile;
    B.march = march ? march : 8;
    B.refill_min = refill_min ? refill_min : 16;
    return GC_OK;
}

int gc_batch_begin(gc_batch *b, const gc_view *views, const float *const *d_tfs, const uint32_t *tf_sizes,
                   uint32_t n_views)
{
    if (!b || !views || !d_tfs || !tf_sizes) return fail(GC_ERR_ARG, "null argument");
    Batch &B = b->impl;
    if (n_views == 0 || n_views > B.max_views) return fail(GC_ERR_ARG, "view count outside 1..max_views");
    uint64_t total = 0;
    for (uint32_t i = 0; i < n_views; ++i) total += (uint64_t)views[i].width * views[i].height;
    if (total > B.max_pixels) return fail(GC_ERR_ARG, "views exceed the batch's max_pixels");
    GC_TRY(gc_batch_set_launch(b, views[0].tile, views[0].march, views[0].refill_min));
    GC_HIP(hipSetDevice(B.C->device));
    for (auto &S : B.slots) S = Batch::Slot{};
    B.table = 0;
    B.cur = -1;
    uint64_t off = 0;
    for (uint32_t i = 0; i < n_views; ++i) {
        GC_TRY(batch_place(B, i, views[i], d_tfs[i], tf_sizes[i], off));
        off += B.slots[i].npix;
    }
    GC_HIP(launch_fill32(reinterpret_cast<uint32_t *>(B.d_stats), sizeof(gc_render_stats) / 4, 0, B.C->stream));
    return GC_OK;
}

int gc_batch_set_view(gc_batch *b, uint32_t slot, const gc_view *view, const float *d_tf, uint32_t tf_size)
{
    if (!b || !view) return fail(GC_ERR_ARG, "null argument");
    Batch &B = b->impl;
    if (slot >= B.max_views) return fail(GC_ERR_ARG, "no such slot");
    if ((uint64_t)view->width * view->height > B.slot_pixels())
        return fail(GC_ERR_ARG, "view larger than a slot (max_pixels / max_views)");
    GC_HIP(hipSetDevice(B.C->device));
    bool busy = false;
    GC_TRY(slot_busy(B, slot, busy));
    if (busy) return fail(GC_ERR_STATE, "slot still has rays waiting for bricks");
    return batch_place(B, slot, *view, d_tf, tf_size, (uint64_t)slot * B.slot_pixels());
}

int gc_batch_clear_view(gc_batch *b, uint32_t slot)
{
    if (!b) return fail(GC_ERR_ARG, "null batch");
    Batch &B = b->impl;
    if (slot >= B.max_views) return fail(GC_ERR_ARG, "no such slot");
    bool busy = false;
    GC_TRY(slot_busy(B, slot, busy));
    if (busy) return fail(GC_ERR_STATE, "slot still has rays waiting for bricks");
    B.slots[slot].used = false;
    B.slots[slot].starting = false;
    while (B.table > 0 && !B.slots[B.table - 1].used) --B.table;
    return GC_OK;
}

int gc_batch_launch(gc_batch *b, uint32_t flags, double *kernel_ms)
{
    if (!b) return fail(GC_ERR_ARG, "null batch");
    if (flags & ~(uint32_t)GC_LAUNCH_STARTING_ONLY) return fail(GC_ERR_ARG, "unknown launch flag");
    Batch &B = b->impl;
    GC_HIP(hipSetDevice(B.C->device));
    GC_TRY(batch_launch(B, nullptr, (flags & GC_LAUNCH_STARTING_ONLY) != 0));
    if (kernel_ms) GC_TRY(batch_kernel_ms(B, *kernel_ms));
    return GC_OK;
}

int gc_batch_pending(gc_batch *b, uint32_t *stopped)
{
    if (!b || !stopped) return fail(GC_ERR_ARG, "null argument");
    Batch &B = b->impl;
    GC_HIP(hipEventSynchronize(B.launched));
    std::memcpy(stopped, B.h_stopped, (size_t)B.max_views * 4);
    for (uint32_t i = 0; i < B.max_views; ++i)
        if (i >= B.table || !B.slots[i].used) stopped[i] = 0;
        else if (B.slots[i].starting) stopped[i] = 0xffffffffu;   // not launched yet
    return GC_OK;
}

int gc_batch_stats(gc_batch *b, gc_render_stats *out)
{
    if (!b || !out) return fail(GC_ERR_ARG, "null argument");
    Batch &B = b->impl;
    GC_HIP(hipMemcpyAsync(out, B.d_stats, sizeof(gc_render_stats), hipMemcpyDeviceToHost, B.C->stream));
    GC_HIP(hipStreamSynchronize(B.C->stream));
    return GC_OK;
}

int gc_batch_render(gc_batch *b, const gc_view *views, const float *const *d_tfs, const uint32_t *tf_sizes,
                    uint32_t n_views, uint32_t max_launches, gc_frame_report *report)
{
    if (!b || !report) return fail(GC_ERR_ARG, "null argument");
    GC_TRY(gc_batch_begin(b, views, d_tfs, tf_sizes, n_views));
    Batch &B = b->impl;
    Cache &C = *B.C;
    std::memset(report, 0, sizeof(*report));
    for (uint32_t it = 0; it < max_launches; ++it) {
        GC_TRY(batch_launch(B, report));
        int done = 0;
        const auto t0 = std::chrono::steady_clock::now();
        GC_TRY(C.update(&done));
        ++report->updates;
        double ms = 0.0;
        GC_TRY(batch_kernel_ms(B, ms));
        report->kernel_ms += ms;
        if (done) {
            report->update_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            report->complete = 1;
            break;
        }
        GC_TRY(gc_cache_wait_idle(B.h));
        const uint32_t before = C.handler.nne;
        GC_TRY(C.update(&done));
        ++report->updates;
        report->bricks_loaded += before;
        report->update_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    GC_TRY(gc_batch_stats(b, &report->stats));
    return GC_OK;
}

int gc_batch_read_pixels(gc_batch *b, uint32_t slot, uint32_t *host)
{
    if (!b || !host) return fail(GC_ERR_ARG, "null argument");
    Batch &B = b->impl;
    if (slot >= B.max_views || !B.slots[slot].used) return fail(GC_ERR_ARG, "no view in this slot");
    GC_HIP(hipSetDevice(B.C->device));
    GC_HIP(hipStreamSynchronize(B.C->stream));
    GC_HIP(hipMemcpy(host, B.slots[slot].v.pixels, B.slots[slot].npix * 4, hipMemcpyDeviceToHost));
    return GC_OK;
}

int gc_batch_read_state(gc_batch *b, uint32_t slot, float *tbuf, void *cstate)
{
    if (!b) return fail(GC_ERR_ARG, "null argument");
    Batch &B = b->impl;
    if (slot >= B.max_views || !B.slots[slot].used) return fail(GC_ERR_ARG, "no view in this slot");
    GC_HIP(hipSetDevice(B.C->device));
    GC_HIP(hipStreamSynchronize(B.C->stream));
    const Batch::Slot &S = B.slots[slot];
    if (tbuf) GC_HIP(hipMemcpy(tbuf, S.v.tbuf, S.npix * 4, hipMemcpyDeviceToHost));
    if (cstate) GC_HIP(hipMemcpy(cstate, S.v.cstate, S.npix * B.cbytes_per_pixel(), hipMemcpyDeviceToHost));
    return GC_OK;
}

int gc_tf_samples(const float *knots, uint32_t nb_knots, uint32_t n, float *out)
{
    if ((!knots && nb_knots) || !out) return fail(GC_ERR_ARG, "null argument");
    tf_samples(knots, nb_knots, n, out);
    return GC_OK;
}

int gc_tf_seeded(uint64_t seed, uint32_t nb_knots, float *knots_out)
{
    if (!knots_out) return fail(GC_ERR_ARG, "null argument");
    tf_seeded(seed, nb_knots, knots_out);
    return GC_OK;
}

int gc_camera_matrix(float rx_deg, float ry_deg, const float t[3], float out12[12])
{
    if (!t || !out12) return fail(GC_ERR_ARG, "null argument");
    camera_matrix(rx_deg, ry_deg, t, out12);
    return GC_OK;
}

} // extern "C"
