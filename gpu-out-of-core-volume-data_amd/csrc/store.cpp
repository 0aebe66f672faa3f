// store.cpp -- the on-disk brick store: the reference's preprocessed volume
// layout (one LZ4 block per big brick under bricks_B_B_B/L<l>/, plus
// bricks.xml), its writer and a brick source that reads it through a host
// LRU (BricksManager.cpp:29-227, MetaData.cpp:36-216, BrickKey.cpp:15-25,
// Bricker_v2.cpp:79-252).
#include "gc_internal.hpp"
#include "source.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <functional>
#include <list>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace gc {

uint64_t lz4_bound(uint64_t n);
int64_t lz4_compress(const uint8_t *src, uint64_t n, uint8_t *dst, uint64_t cap);
int64_t lz4_decompress(const uint8_t *src, uint64_t n, uint8_t *dst, uint64_t dst_n);

namespace {

// Bkey: uint16_t[4] = {level, x, y, z} read as one little-endian uint64
uint64_t bkey(uint32_t level, uint32_t x, uint32_t y, uint32_t z)
{
    return (uint64_t)(level & 0xffffu) | (uint64_t)(x & 0xffffu) << 16 | (uint64_t)(y & 0xffffu) << 32 |
           (uint64_t)(z & 0xffffu) << 48;
}

// The reference appends the folder to the volume directory as is (its
// directories end in '/'); a directory given without one gets it.
std::string dir_of(const char *volume_dir)
{
    std::string d(volume_dir ? volume_dir : "");
    if (!d.empty() && d.back() != '/') d.push_back('/');
    return d;
}
std::string folder_of(uint32_t bs)
{
    const std::string b = std::to_string(bs);
    return "bricks_" + b + "_" + b + "_" + b;
}
std::string brick_path(const std::string &dir, uint32_t bs, uint32_t level, const uint32_t xyz[3])
{
    const std::string l = std::to_string(level);
    return dir + folder_of(bs) + "/L" + l + "/L" + l + "_" + std::to_string(xyz[0]) + "_" + std::to_string(xyz[1]) +
           "_" + std::to_string(xyz[2]) + ".raw";
}

bool make_dir(const std::string &p)
{
    if (::mkdir(p.c_str(), 0755) == 0 || errno == EEXIST) return true;
    return false;
}
// every missing directory of `p` (a directory path)
bool make_dirs(const std::string &p)
{
    for (size_t i = 1; i <= p.size(); ++i)
        if (i == p.size() || p[i] == '/')
            if (!make_dir(p.substr(0, i))) return false;
    return true;
}

bool read_file(const std::string &path, std::vector<uint8_t> &out)
{
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    out.clear();
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + n);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

bool write_file(const std::string &path, const uint8_t *p, size_t n)
{
    std::FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(p, 1, n, f) == n;
    return std::fclose(f) == 0 && ok;
}

// Big bricks per axis of a level: the level's real size over the inner
// brick size, over BigBrickSize (Bricker_v2.cpp:50-74).
void big_bricks(const gc_volume_cfg &vc, uint32_t level, uint32_t out[3])
{
    for (int a = 0; a < 3; ++a) {
        const uint32_t inner = vc.brick_size[a] - 2 * vc.covering[a];
        out[a] = vc.real_sizes[level][a] / inner / vc.big_brick_size[a];
    }
}

// Attribute values of every <tag .../> inside <section>...</section>.
std::vector<std::string> xml_values(const std::string &text, const char *section, const char *tag, const char *attr)
{
    std::vector<std::string> out;
    const std::string open = std::string("<") + section, close = std::string("</") + section + ">";
    const size_t s = text.find(open);
    if (s == std::string::npos) return out;
    const size_t e = text.find(close, s);
    const std::string t = std::string("<") + tag, key = std::string(attr) + "=\"";
    for (size_t p = text.find(t, s); p != std::string::npos && (e == std::string::npos || p < e);
         p = text.find(t, p + 1)) {
        const char c = p + t.size() < text.size() ? text[p + t.size()] : '\0';
        if (c != ' ' && c != '/' && c != '>') continue;   // a longer tag name
        const size_t end = text.find('>', p);
        const size_t k = text.find(key, p);
        if (k == std::string::npos || k > end) continue;
        const size_t v0 = k + key.size(), v1 = text.find('"', v0);
        out.push_back(text.substr(v0, v1 - v0));
    }
    return out;
}

// MetaData::write_bricks_xml (MetaData.cpp:36-103), in tinyxml2's layout.
bool write_bricks_xml(const std::string &path, const gc_volume_cfg &vc, const std::vector<uint64_t> &empty,
                      const std::vector<float> &histo)
{
    std::string x = "<Meta>\n    <Levels>\n";
    for (uint32_t l = 0; l < vc.nb_levels; ++l) {
        char line[512];
        std::snprintf(line, sizeof(line),
                      "        <Level level=\"%u\" initial_size_X=\"%u\" initial_size_Y=\"%u\" initial_size_Z=\"%u\" "
                      "real_size_X=\"%u\" real_size_Y=\"%u\" real_size_Z=\"%u\"/>\n",
                      l, vc.initial_sizes[l][0], vc.initial_sizes[l][1], vc.initial_sizes[l][2], vc.real_sizes[l][0],
                      vc.real_sizes[l][1], vc.real_sizes[l][2]);
        x += line;
    }
    x += "    </Levels>\n";
    if (!empty.empty()) {
        x += "    <EmptyBricks>\n";
        for (uint64_t k : empty) x += "        <Brick ID=\"" + std::to_string((int64_t)k) + "\"/>\n";
        x += "    </EmptyBricks>\n";
    }
    x += "    <Histogram>\n";
    for (float h : histo) {
        char v[64];
        std::snprintf(v, sizeof(v), "%.8g", (double)h);
        x += std::string("        <Histo Value=\"") + v + "\"/>\n";
    }
    x += "    </Histogram>\n</Meta>\n";
    return write_file(path, reinterpret_cast<const uint8_t *>(x.data()), x.size());
}

bool read_meta(const std::string &xml_path, std::vector<uint64_t> &empty, std::vector<float> &histo)
{
    std::vector<uint8_t> raw;
    if (!read_file(xml_path, raw)) return false;
    const std::string text(raw.begin(), raw.end());
    if (text.find("<Levels") == std::string::npos) return false;
    empty.clear();
    histo.clear();
    for (const std::string &v : xml_values(text, "EmptyBricks", "Brick", "ID"))
        empty.push_back((uint64_t)std::strtoull(v.c_str(), nullptr, 10));
    for (const std::string &v : xml_values(text, "Histogram", "Histo", "Value"))
        histo.push_back(std::strtof(v.c_str(), nullptr));
    return true;
}

void run_threads(uint32_t n, uint32_t threads, const std::function<void(uint32_t)> &fn)
{
    std::atomic<uint32_t> next{0};
    auto work = [&] {
        for (uint32_t i; (i = next.fetch_add(1)) < n;) fn(i);
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < std::min(threads, n); ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
}

} // namespace

// Big bricks read from the store through a host LRU (BricksManager).
struct StoreSource : Source {
    std::string dir;
    std::unordered_set<uint64_t> empty;
    uint32_t cap = 32768;
    std::mutex m;
    std::list<uint64_t> lru;   // front: most recently used
    std::unordered_map<uint64_t, std::pair<std::vector<uint8_t>, std::list<uint64_t>::iterator>> held;
    std::atomic<uint64_t> hits{0}, loads{0}, empties{0};
    uint32_t threads() const override
    {
        const uint32_t h = std::thread::hardware_concurrency();
        return h ? h : 1;
    }
    int get(uint32_t level, const uint32_t bbc[3], void *dst) override
    {
        if (level >= vc.nb_levels) return GC_BRICK_ERROR;
        const uint64_t key = bkey(level, bbc[0], bbc[1], bbc[2]);
        if (empty.count(key)) {
            ++empties;
            return GC_BRICK_EMPTY;
        }
        {
            std::lock_guard<std::mutex> lk(m);
            auto it = held.find(key);
            if (it != held.end()) {
                std::memcpy(dst, it->second.first.data(), bb_bytes);
                lru.splice(lru.begin(), lru, it->second.second);
                ++hits;
                return GC_BRICK_SUCCESS;
            }
        }
        std::vector<uint8_t> file;
        if (!read_file(brick_path(dir, bs, level, bbc), file)) return GC_BRICK_ERROR;
        uint8_t *out = static_cast<uint8_t *>(dst);
        if (lz4_decompress(file.data(), file.size(), out, bb_bytes) < 0) {
            // an uncompressed brick (write_brick without compression)
            if (file.size() != bb_bytes) return GC_BRICK_ERROR;
            std::memcpy(out, file.data(), bb_bytes);
        }
        ++loads;
        if (cap == 0) return GC_BRICK_SUCCESS;
        std::lock_guard<std::mutex> lk(m);
        if (!held.count(key)) {
            lru.push_front(key);
            auto &e = held[key];
            e.first.assign(out, out + bb_bytes);
            e.second = lru.begin();
            while (held.size() > cap) {
                held.erase(lru.back());
                lru.pop_back();
            }
        }
        return GC_BRICK_SUCCESS;
    }
};

} // namespace gc

using namespace gc;

extern "C" uint64_t gc_brick_key(uint32_t level, uint32_t x, uint32_t y, uint32_t z) { return bkey(level, x, y, z); }

extern "C" void gc_brick_key_unpack(uint64_t key, uint32_t *level, uint32_t xyz[3])
{
    if (level) *level = (uint32_t)(key & 0xffffu);
    if (xyz) {
        xyz[0] = (uint32_t)(key >> 16 & 0xffffu);
        xyz[1] = (uint32_t)(key >> 32 & 0xffffu);
        xyz[2] = (uint32_t)(key >> 48 & 0xffffu);
    }
}

extern "C" int gc_store_brick_path(const char *volume_dir, uint32_t brick_size, uint32_t level, const uint32_t xyz[3],
                                   char *out, size_t out_size)
{
    if (!volume_dir || !xyz || !out || out_size == 0) return fail(GC_ERR_ARG, "null argument");
    const std::string p = brick_path(dir_of(volume_dir), brick_size, level, xyz);
    if (p.size() + 1 > out_size) return fail(GC_ERR_ARG, "path buffer too small");
    std::memcpy(out, p.c_str(), p.size() + 1);
    return GC_OK;
}

extern "C" int gc_store_write(const gc_volume_cfg *vc, gc_source *src, const char *volume_dir, int compression,
                              float empty_threshold, const float *histogram, uint32_t histo_n, uint32_t threads)
{
    if (!vc || !src || !volume_dir) return fail(GC_ERR_ARG, "null argument");
    Source &S = *src->impl;
    if (S.bs != vc->brick_size[0] || S.bytes != vc->encoded_bytes || S.cov != vc->covering[0] ||
        S.BB[0] != vc->big_brick_size[0] || S.BB[1] != vc->big_brick_size[1] || S.BB[2] != vc->big_brick_size[2])
        return fail(GC_ERR_ARG, "gc_store_write: the source's brick layout differs from the volume configuration");
    if (histogram == nullptr && histo_n != 0) return fail(GC_ERR_ARG, "histo_n without a histogram");
    const std::string dir = dir_of(volume_dir), root = dir + folder_of(S.bs);
    if (!make_dirs(root)) return fail(GC_ERR_SOURCE, "cannot create " + root);
    std::vector<std::array<uint32_t, 4>> jobs;   // level, x, y, z
    for (uint32_t l = 0; l < vc->nb_levels; ++l) {
        if (!make_dir(root + "/L" + std::to_string(l))) return fail(GC_ERR_SOURCE, "cannot create level directory");
        uint32_t nb[3];
        big_bricks(*vc, l, nb);
        for (uint32_t z = 0; z < nb[2]; ++z)
            for (uint32_t y = 0; y < nb[1]; ++y)
                for (uint32_t x = 0; x < nb[0]; ++x) jobs.push_back({l, x, y, z});
    }
    std::mutex em;
    std::vector<uint64_t> empty;
    std::atomic<int> err{GC_OK};
    std::string why;
    const uint32_t nth = S.thread_safe() ? (threads ? threads : std::max(1u, std::thread::hardware_concurrency())) : 1u;
    run_threads((uint32_t)jobs.size(), nth, [&](uint32_t i) {
        if (err.load() != GC_OK) return;
        const auto &j = jobs[i];
        const uint32_t bbc[3] = {j[1], j[2], j[3]};
        std::vector<uint8_t> brick(S.bb_bytes, 0), packed;
        const int st = S.get(j[0], bbc, brick.data());
        if (st == GC_BRICK_ERROR) {
            err = GC_ERR_SOURCE;
            return;
        }
        bool is_empty = st == GC_BRICK_EMPTY;
        if (is_empty) std::fill(brick.begin(), brick.end(), 0);
        else if (empty_threshold > 0.f) {
            // BrickProcessorPredicate.hpp:11-45 over the whole big brick
            is_empty = true;
            for (uint64_t v = 0; v < S.bb_vox && is_empty; ++v) {
                float x;
                if (S.bytes == 1) x = (float)brick[v];
                else if (S.bytes == 2) x = (float)reinterpret_cast<const uint16_t *>(brick.data())[v];
                else x = reinterpret_cast<const float *>(brick.data())[v];
                if (!(x < empty_threshold)) is_empty = false;
            }
        }
        if (is_empty) {
            std::lock_guard<std::mutex> lk(em);
            empty.push_back(bkey(j[0], j[1], j[2], j[3]));
        }
        const uint8_t *out = brick.data();
        size_t n = brick.size();
        if (compression && S.bb_bytes < 0x7E000000ull) {
            packed.resize(lz4_bound(S.bb_bytes));
            const int64_t c = lz4_compress(brick.data(), S.bb_bytes, packed.data(), packed.size());
            if (c < 0) {
                err = GC_ERR_SOURCE;
                return;
            }
            out = packed.data();
            n = (size_t)c;
        }
        if (!write_file(brick_path(dir, S.bs, j[0], bbc), out, n)) err = GC_ERR_SOURCE;
    });
    if (err.load() != GC_OK) return fail(err.load(), "gc_store_write: a brick could not be read from the source or written");
    std::sort(empty.begin(), empty.end());
    std::vector<float> histo;
    if (histogram) histo.assign(histogram, histogram + histo_n);
    else if (vc->encoded_bytes == 1) histo.assign(256, 0.f);
    else if (vc->encoded_bytes == 2) histo.assign(65535, 0.f);
    if (!write_bricks_xml(root + "/bricks.xml", *vc, empty, histo)) return fail(GC_ERR_SOURCE, "cannot write bricks.xml");
    return GC_OK;
}

extern "C" int gc_store_read_meta(const char *volume_dir, uint32_t brick_size, uint64_t *empty_keys, uint32_t *n_empty,
                                  float *histogram, uint32_t *histo_n)
{
    if (!volume_dir) return fail(GC_ERR_ARG, "null argument");
    std::vector<uint64_t> empty;
    std::vector<float> histo;
    const std::string xml = dir_of(volume_dir) + folder_of(brick_size) + "/bricks.xml";
    if (!read_meta(xml, empty, histo)) return fail(GC_ERR_SOURCE, "cannot read " + xml);
    if (n_empty) {
        if (empty_keys) std::memcpy(empty_keys, empty.data(), std::min<size_t>(*n_empty, empty.size()) * 8);
        *n_empty = (uint32_t)empty.size();
    }
    if (histo_n) {
        if (histogram) std::memcpy(histogram, histo.data(), std::min<size_t>(*histo_n, histo.size()) * 4);
        *histo_n = (uint32_t)histo.size();
    }
    return GC_OK;
}

extern "C" int gc_source_create_store(const gc_volume_cfg *vc, const char *volume_dir, uint32_t cache_bricks,
                                      gc_source **out)
{
    if (!vc || !volume_dir || !out) return fail(GC_ERR_ARG, "null argument");
    auto s = std::make_unique<StoreSource>();
    s->init(*vc);
    s->dir = dir_of(volume_dir);
    s->cap = cache_bricks ? cache_bricks : 32768u;
    std::vector<uint64_t> empty;
    std::vector<float> histo;
    const std::string xml = s->dir + folder_of(s->bs) + "/bricks.xml";
    if (!read_meta(xml, empty, histo)) return fail(GC_ERR_SOURCE, "cannot read " + xml);
    s->empty.insert(empty.begin(), empty.end());
    *out = new gc_source{std::move(s)};
    return GC_OK;
}

extern "C" int gc_store_stats(gc_source *src, uint64_t out[4])
{
    if (!src || !out) return fail(GC_ERR_ARG, "null argument");
    auto *s = dynamic_cast<StoreSource *>(src->impl.get());
    if (!s) return fail(GC_ERR_ARG, "not a store source");
    std::lock_guard<std::mutex> lk(s->m);
    out[0] = s->hits.load();
    out[1] = s->loads.load();
    out[2] = s->held.size();
    out[3] = s->empties.load();
    return GC_OK;
}
