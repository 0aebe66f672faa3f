// server.cpp -- the serving loop (gc_server_*): continuous batching of view
// requests over several resident volumes, admission by brick-pool memory.
//
// One scheduler thread owns every device call.  Each iteration it
//   1. admits queued requests: a volume's cache (page directory + pool) and
//      view batch are created when its first request arrives and the pools
//      of the resident volumes plus its own fit the budget, evicting volumes
//      with nothing in flight (least recently used first) to make room; free
//      slots of a resident volume take its queued requests in arrival order;
//   2. visits every resident volume without waiting for the others: when the
//      volume's launch (on the volume's own stream) has finished, it updates
//      the cache (CacheManager::update), reads the per-slot miss counters and
//      retires the views with no ray left waiting (pixels copied to the host,
//      latency recorded); a volume with no launch in flight sets its newly
//      admitted views and launches again if it has rays to trace.
// A volume whose rays all wait for the request handler is not relaunched
// until the handler is done; the update that then loads the bricks runs
// before the next launch.  The bricks and the LRU of each volume evolve by
// the reference's update rules; only the interleaving of views differs.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gc_internal.hpp"

namespace gc {

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    uint32_t volume = 0;
    gc_view view{};
    void *d_tf = nullptr;
    uint32_t tf_size = 0;
    double t_submit = 0.0, t_done = 0.0;
    uint32_t launches = 0;
    std::vector<uint32_t> pixels;
};

struct Volume {
    gc_volume_cfg vc{};
    gc_source *src = nullptr;
    uint64_t pool_bytes = 0, footprint = 0;
    uint32_t max_requests = 0;
    gc_cache *cache = nullptr;
    gc_batch *batch = nullptr;
    std::deque<Request *> queue;
    std::vector<Request *> slots;
    std::vector<bool> fresh;            // slot placed, not launched yet
    std::vector<bool> unplaced;         // slot taken, its view not yet set in the batch
    std::vector<std::pair<uint32_t, Request *>> pending;   // (slot, request) to set before the next launch
    bool load_pending = false;
    bool in_flight = false;             // a launch is running; `done` is recorded after it
    hipEvent_t done = nullptr;
    uint64_t last_use = 0, samples_seen = 0, bricks_seen = 0, starts_seen = 0;
    double fetch_seen = 0.0;
    // concurrent views admitted (limit by pool pressure); pressure seen
    // since the last view of this volume completed
    uint32_t cap = 0;
    bool pressure = false;
    uint32_t active() const
    {
        uint32_t k = 0;
        for (auto *r : slots) k += r ? 1u : 0u;
        return k;
    }
    bool resident() const { return cache != nullptr; }
    bool busy_slots() const
    {
        for (auto *r : slots)
            if (r) return true;
        return false;
    }
};

} // namespace

struct Server {
    gc_server_cfg cfg{};
    int device = 0;
    // volumes live at stable addresses: the scheduler works on them without
    // the lock while gc_server_add_volume may append to the list
    std::vector<std::unique_ptr<Volume>> vols;
    std::mutex m;
    std::condition_variable cv, done_cv;
    std::thread worker;
    bool stop = false;
    std::string error;
    std::map<uint64_t, Request *> reqs;
    std::deque<Request *> finished;     // completed, not yet polled
    uint64_t next_id = 1, in_flight = 0, tick = 0;
    uint64_t resident_bytes = 0;
    gc_server_stats st{};
    Clock::time_point t0 = Clock::now();

    double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
    uint64_t slot_bytes() const
    {
        // pixels + t + colour state + resume position + opacity product + two
        // ray-list entries + the stop flag per pixel
        return (uint64_t)cfg.max_views * cfg.slot_pixels * (4 + 4 + (cfg.precision == 64 ? 32 : 16) + 16 + 8 + 8 + 1);
    }
    struct Place {
        Volume *v;
        uint32_t slot;
        Request *r;
    };
    void loop();
    int iterate(std::unique_lock<std::mutex> &lk);
    int device_phase(const std::vector<Volume *> &all, const std::vector<Volume *> &evict,
                     const std::vector<Volume *> &create, const std::vector<Place> &place,
                     std::vector<Request *> &done, gc_server_stats &d, bool &progressed);
    bool progressed = false;
    int make_resident(Volume &v);
    void release(Volume &v);
};

void Server::release(Volume &v)
{
    if (v.done) hipEventSynchronize(v.done);
    if (v.batch) gc_batch_destroy(v.batch);
    if (v.cache) gc_cache_destroy(v.cache);
    if (v.done) hipEventDestroy(v.done);
    v.batch = nullptr;
    v.cache = nullptr;
    v.done = nullptr;
    v.in_flight = false;
    v.load_pending = false;
    v.samples_seen = 0;
    v.starts_seen = 0;
    v.bricks_seen = 0;
    v.fetch_seen = 0.0;
}

int Server::make_resident(Volume &v)
{
    gc_cache_cfg cc{};
    const uint32_t bs = v.vc.brick_size[0];
    const uint64_t brick_bytes = (uint64_t)bs * bs * bs * v.vc.encoded_bytes;
    GC_TRY(gc_pool_shape(v.pool_bytes, brick_bytes, cc.cache_size));
    for (int a = 0; a < 3; ++a) cc.block_size[a] = bs;
    cc.data_cache_flags = 1;
    cc.max_requests = v.max_requests;
    cc.sync_requests = 0;
    GC_TRY(gc_cache_create(&v.vc, &cc, v.src, device, &v.cache));
    const int rc = gc_batch_create(v.cache, cfg.max_views, (uint64_t)cfg.max_views * cfg.slot_pixels, cfg.precision,
                                   &v.batch);
    if (rc != GC_OK) {
        gc_cache_destroy(v.cache);
        v.cache = nullptr;
        return rc;
    }
    GC_TRY(gc_batch_set_launch(v.batch, cfg.tile, cfg.march, 0));
    GC_TRY(gc_batch_set_graph(v.batch, cfg.graph ? 1 : 0));
    GC_HIP(hipEventCreateWithFlags(&v.done, hipEventDisableTiming));
    return GC_OK;
}

// Device work of one iteration (no lock held): evictions, new pools, then
// per volume, without waiting for the others: a finished launch is followed
// by its cache update and the retirement of completed views; a volume with
// no launch in flight takes its new views and launches again.
int Server::device_phase(const std::vector<Volume *> &all, const std::vector<Volume *> &evict,
                         const std::vector<Volume *> &create, const std::vector<Place> &place,
                         std::vector<Request *> &done, gc_server_stats &d, bool &progressed)
{
    progressed = !evict.empty() || !create.empty();
    for (Volume *v : evict) release(*v);
    for (Volume *v : create) GC_TRY(make_resident(*v));
    for (const Place &p : place) p.v->pending.emplace_back(p.slot, p.r);
    std::vector<uint32_t> pend(cfg.max_views);
    for (Volume *vp : all) {
        Volume &v = *vp;
        if (!v.resident()) continue;
        if (v.in_flight) {
            const hipError_t q = hipEventQuery(v.done);
            if (q == hipErrorNotReady) continue;   // still running
            GC_HIP(q);
            v.in_flight = false;
            progressed = true;
            // update and retirement
            const auto t1 = Clock::now();
            int complete = 0;
            gc_cache_counters before{}, after{};
            GC_TRY(gc_cache_counters_get(v.cache, &before));
            GC_TRY(gc_cache_update(v.cache, &complete));
            GC_TRY(gc_cache_counters_get(v.cache, &after));
            if (!complete) v.load_pending = true;
            if (after.requests != before.requests &&
                (uint64_t)after.last_used + after.last_requested > after.slots) {
                // the working set of the views in flight no longer fits the pool:
                // admit no more than half of them until views complete
                ++d.pressure_events;
                v.pressure = true;
                v.cap = std::max(1u, v.active() / 2);
            }
            d.bricks_loaded += after.bricks_loaded - v.bricks_seen;
            v.bricks_seen = after.bricks_loaded;
            d.fetch_ms += after.fetch_ms - v.fetch_seen;
            v.fetch_seen = after.fetch_ms;
            GC_TRY(gc_batch_pending(v.batch, pend.data()));
            for (uint32_t s = 0; s < cfg.max_views; ++s) {
                Request *r = v.slots[s];
                if (!r || v.unplaced[s]) continue;   // taken after the launch was issued
                ++r->launches;
                v.fresh[s] = false;
                if (pend[s] != 0) continue;
                r->pixels.resize((size_t)r->view.width * r->view.height);
                GC_TRY(gc_batch_read_pixels(v.batch, s, r->pixels.data()));
                GC_TRY(gc_batch_clear_view(v.batch, s));
                r->t_done = now_ms();
                v.slots[s] = nullptr;
                done.push_back(r);
                if (!v.pressure) v.cap = std::min(cfg.max_views, v.cap + 1);
                v.pressure = false;
            }
            gc_render_stats rs{};
            GC_TRY(gc_batch_stats(v.batch, &rs));
            d.samples += rs.mapped_samples - v.samples_seen;
            v.samples_seen = rs.mapped_samples;
            d.ray_starts += rs.pixels - v.starts_seen;
            v.starts_seen = rs.pixels;
            v.last_use = ++tick;
            ++d.launches;
            d.update_ms += std::chrono::duration<double, std::milli>(Clock::now() - t1).count();
        }
        // new views join before the next launch
        const auto t0 = Clock::now();
        for (const auto &ps : v.pending) {
            Request *r = ps.second;
            GC_TRY(gc_batch_set_view(v.batch, ps.first, &r->view, static_cast<const float *>(r->d_tf), r->tf_size));
            v.unplaced[ps.first] = false;
            progressed = true;
        }
        v.pending.clear();
        if (!v.busy_slots()) continue;
        bool fresh = false;
        for (uint32_t s = 0; s < cfg.max_views; ++s) fresh = fresh || (v.slots[s] && v.fresh[s]);
        int busy = 0;
        GC_TRY(gc_cache_busy(v.cache, &busy));
        if (busy && !fresh) continue;                   // every ray waits for the handler
        if (!busy && v.load_pending) {
            int complete = 0;
            GC_TRY(gc_cache_update(v.cache, &complete));   // load what the handler fetched
            v.load_pending = false;
        }
        // while the handler fetches, only the starting views are traced
        GC_TRY(gc_batch_launch(v.batch, busy ? GC_LAUNCH_STARTING_ONLY : 0u, nullptr));
        void *stream = nullptr;
        GC_TRY(gc_cache_stream(v.cache, &stream));
        GC_HIP(hipEventRecord(v.done, static_cast<hipStream_t>(stream)));
        v.in_flight = true;
        progressed = true;
        d.launch_ms += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    }
    if (!progressed) ++d.idle_iterations;
    return GC_OK;
}

// One scheduler iteration; called and returns with the lock held.
int Server::iterate(std::unique_lock<std::mutex> &lk)
{
    // admission decisions
    std::vector<Volume *> evict, create;
    std::vector<Place> place;
    uint64_t projected = resident_bytes;
    std::vector<Volume *> all;
    for (auto &vp : vols) all.push_back(vp.get());
    for (Volume *vp : all) {
        Volume &v = *vp;
        if (v.queue.empty()) continue;
        if (cfg.plain && in_flight > 0) break;
        // a volume picked for eviction in this pass is released before any
        // placement takes effect, so it counts as not resident here
        const bool evicted = std::find(evict.begin(), evict.end(), &v) != evict.end();
        const bool now_resident = (v.resident() && !evicted) ||
                                  std::find(create.begin(), create.end(), &v) != create.end();
        if (!now_resident) {
            while (projected + v.footprint > cfg.pool_budget) {
                // victims: resident, nothing in flight and nothing queued (a
                // volume with queued requests is about to be used again)
                Volume *victim = nullptr;
                for (Volume *u : all)
                    if (u != &v && u->resident() && !u->busy_slots() && u->queue.empty() &&
                        std::find(evict.begin(), evict.end(), u) == evict.end() &&
                        (!victim || u->last_use < victim->last_use))
                        victim = u;
                if (!victim) break;
                evict.push_back(victim);
                projected -= victim->footprint;
            }
            if (projected + v.footprint > cfg.pool_budget) continue;   // waits for memory
            create.push_back(&v);
            projected += v.footprint;
        }
        if (v.cap == 0 || cfg.fixed_views) v.cap = cfg.max_views;
        uint32_t active = v.active();
        for (uint32_t s = 0; s < cfg.max_views && !v.queue.empty(); ++s) {
            if (v.slots[s]) continue;
            if (cfg.plain && in_flight > 0) break;
            if (active >= v.cap) break;   // pool pressure: wait for views to complete
            ++active;
            Request *r = v.queue.front();
            v.queue.pop_front();
            v.slots[s] = r;
            v.fresh[s] = true;
            v.unplaced[s] = true;
            ++in_flight;
            place.push_back({&v, s, r});
        }
    }
    resident_bytes = projected;
    st.pools_evicted += evict.size();
    st.pools_created += create.size();
    lk.unlock();
    std::vector<Request *> done;
    gc_server_stats d{};
    const int rc = device_phase(all, evict, create, place, done, d, progressed);
    lk.lock();
    st.samples += d.samples;
    st.ray_starts += d.ray_starts;
    st.launches += d.launches;
    st.bricks_loaded += d.bricks_loaded;
    st.pressure_events += d.pressure_events;
    ++st.iterations;
    st.idle_iterations += d.idle_iterations;
    st.launch_ms += d.launch_ms;
    st.update_ms += d.update_ms;
    st.fetch_ms += d.fetch_ms;
    for (Request *r : done) {
        if (r->d_tf) gc_free(r->d_tf);
        r->d_tf = nullptr;
        finished.push_back(r);
        --in_flight;
        ++st.completed;
    }
    if (!done.empty()) done_cv.notify_all();
    return rc;
}

void Server::loop()
{
    hipSetDevice(device);
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
        cv.wait(lk, [&] {
            if (stop) return true;
            for (auto &v : vols)
                if (!v->queue.empty() || v->busy_slots()) return true;
            return false;
        });
        if (stop) break;
        const int rc = iterate(lk);
        if (rc != GC_OK) {
            error = gc_last_error();
            done_cv.notify_all();
            break;
        }
        if (!progressed) {
            // launches still running and handlers still fetching: give the
            // host cores to the handlers for a moment
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            lk.lock();
        }
    }
}

} // namespace gc

struct gc_server {
    gc::Server impl;
};

using namespace gc;

extern "C" {

int gc_server_create(const gc_server_cfg *cfg, int device, gc_server **out)
{
    if (!cfg || !out) return fail(GC_ERR_ARG, "null argument");
    if (cfg->max_views == 0 || cfg->max_views > DVR_MAX_VIEWS) return fail(GC_ERR_ARG, "max_views must be 1..64");
    if (cfg->slot_pixels == 0) return fail(GC_ERR_ARG, "slot_pixels must be positive");
    if (cfg->precision != 32 && cfg->precision != 64) return fail(GC_ERR_ARG, "precision must be 32 or 64");
    if (cfg->tile >= GC_TILES || cfg->march > 8) return fail(GC_ERR_ARG, "bad launch shape");
    auto s = std::make_unique<gc_server>();
    s->impl.cfg = *cfg;
    s->impl.device = device;
    GC_HIP(hipSetDevice(device));
    s->impl.worker = std::thread([p = &s->impl] { p->loop(); });
    *out = s.release();
    return GC_OK;
}

int gc_server_destroy(gc_server *s)
{
    if (!s) return GC_OK;
    Server &S = s->impl;
    {
        std::lock_guard<std::mutex> lk(S.m);
        S.stop = true;
    }
    S.cv.notify_all();
    if (S.worker.joinable()) S.worker.join();
    hipSetDevice(S.device);
    for (auto &v : S.vols)
        if (v->resident()) S.release(*v);
    for (auto &kv : S.reqs) {
        if (kv.second->d_tf) gc_free(kv.second->d_tf);
        delete kv.second;
    }
    delete s;
    return GC_OK;
}

int gc_server_add_volume(gc_server *s, const gc_volume_cfg *vc, gc_source *src, uint64_t pool_bytes,
                         uint32_t max_requests, uint32_t *volume)
{
    if (!s || !vc || !src || !volume) return fail(GC_ERR_ARG, "null argument");
    Server &S = s->impl;
    const uint64_t bs = vc->brick_size[0];
    const uint64_t brick_bytes = bs * bs * bs * vc->encoded_bytes;
    if (pool_bytes < brick_bytes) return fail(GC_ERR_ARG, "pool smaller than one brick");
    if (max_requests == 0) return fail(GC_ERR_ARG, "max_requests must be positive");
    std::lock_guard<std::mutex> lk(S.m);
    auto v = std::make_unique<Volume>();
    v->vc = *vc;
    v->src = src;
    v->pool_bytes = pool_bytes;
    v->max_requests = max_requests;
    v->footprint = pool_bytes + S.slot_bytes();
    if (v->footprint > S.cfg.pool_budget) return fail(GC_ERR_ARG, "volume's pool and view slots exceed the budget");
    v->slots.assign(S.cfg.max_views, nullptr);
    v->fresh.assign(S.cfg.max_views, false);
    v->unplaced.assign(S.cfg.max_views, false);
    S.vols.push_back(std::move(v));
    *volume = (uint32_t)(S.vols.size() - 1);
    return GC_OK;
}

int gc_server_submit(gc_server *s, uint32_t volume, const gc_view *view, const float *tf, uint32_t tf_size,
                     uint64_t *id)
{
    if (!s || !view || !tf || !id || tf_size == 0) return fail(GC_ERR_ARG, "null argument");
    Server &S = s->impl;
    if ((uint64_t)view->width * view->height > S.cfg.slot_pixels || view->width == 0 || view->height == 0)
        return fail(GC_ERR_ARG, "view larger than slot_pixels");
    if (view->precision != S.cfg.precision) return fail(GC_ERR_ARG, "view precision differs from the server's");
    if (view->compositing != 0) return fail(GC_ERR_ARG, "the server renders DVR only (compositing 0)");
    auto r = std::make_unique<Request>();
    r->view = *view;
    r->tf_size = tf_size;
    GC_HIP(hipSetDevice(S.device));
    GC_TRY(gc_malloc(&r->d_tf, (size_t)tf_size * 16));
    const int rc = gc_memcpy_h2d(r->d_tf, tf, (size_t)tf_size * 16);
    if (rc != GC_OK) {
        gc_free(r->d_tf);
        return rc;
    }
    {
        std::lock_guard<std::mutex> lk(S.m);
        if (!S.error.empty()) {
            gc_free(r->d_tf);
            return fail(GC_ERR_STATE, "server stopped: " + S.error);
        }
        if (volume >= S.vols.size()) {
            gc_free(r->d_tf);
            return fail(GC_ERR_ARG, "no such volume");
        }
        r->id = S.next_id++;
        r->volume = volume;
        r->t_submit = S.now_ms();
        *id = r->id;
        S.vols[volume]->queue.push_back(r.get());
        S.reqs[r->id] = r.get();
        ++S.st.submitted;
        r.release();
    }
    S.cv.notify_all();
    return GC_OK;
}

int gc_server_poll(gc_server *s, gc_server_result *out, uint32_t max, uint32_t *n)
{
    if (!s || !n || (!out && max)) return fail(GC_ERR_ARG, "null argument");
    Server &S = s->impl;
    std::lock_guard<std::mutex> lk(S.m);
    uint32_t k = 0;
    while (k < max && !S.finished.empty()) {
        Request *r = S.finished.front();
        S.finished.pop_front();
        gc_server_result &o = out[k++];
        o.id = r->id;
        o.volume = r->volume;
        o.width = r->view.width;
        o.height = r->view.height;
        o.launches = r->launches;
        o.submit_ms = r->t_submit;
        o.latency_ms = r->t_done - r->t_submit;
    }
    *n = k;
    if (!S.error.empty() && k == 0) return fail(GC_ERR_STATE, "server stopped: " + S.error);
    return GC_OK;
}

int gc_server_take_pixels(gc_server *s, uint64_t id, uint32_t *host)
{
    if (!s || !host) return fail(GC_ERR_ARG, "null argument");
    Server &S = s->impl;
    std::lock_guard<std::mutex> lk(S.m);
    auto it = S.reqs.find(id);
    if (it == S.reqs.end() || it->second->t_done == 0.0) return fail(GC_ERR_ARG, "request not completed");
    Request *r = it->second;
    std::memcpy(host, r->pixels.data(), r->pixels.size() * 4);
    delete r;
    S.reqs.erase(it);
    return GC_OK;
}

int gc_server_drain(gc_server *s)
{
    if (!s) return fail(GC_ERR_ARG, "null server");
    Server &S = s->impl;
    std::unique_lock<std::mutex> lk(S.m);
    S.done_cv.wait(lk, [&] { return S.st.completed == S.st.submitted || !S.error.empty(); });
    if (!S.error.empty()) return fail(GC_ERR_STATE, "server stopped: " + S.error);
    return GC_OK;
}

int gc_server_stats_get(gc_server *s, gc_server_stats *out)
{
    if (!s || !out) return fail(GC_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->impl.m);
    *out = s->impl.st;
    return GC_OK;
}

} // extern "C"
