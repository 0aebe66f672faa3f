// tdns/gpucache/CacheManager.hpp -- C++ drop-in for GcCore's cache manager
// (GcCore/include/GcCore/cuda/libGPUCache/CacheManager.hpp:38-128) and the
// configuration types it takes (libData/VolumeConfiguration.hpp:14-31,
// libData/CacheConfiguration.hpp:13-19, libData/MetaData.hpp:16-103),
// header-only over the C ABI of libgcvol.so (gcvol.h).
//
// Same names, arguments and error behaviour as the reference: a failed call
// throws std::runtime_error with the library's message.  Differences a caller
// sees:
//  * MetaData carries the brick source the reference reads through its
//    BricksManager (LZ4 brick files).  load_raw_volume() builds the mip
//    pyramid of a raw volume on the GPU in memory (the Mipmapper + Bricker
//    preprocessing, Mipmapper.cu:401-597, Bricker_v2.cpp:45-252) and serves
//    bricks from it; load() opens a preprocessed store (bricks_B_B_B/L*/ LZ4
//    brick files and bricks.xml, read through a host LRU as BricksManager
//    does) with its empty list and histogram; set_source() attaches any
//    gc_source (synthetic, callback).  The MetaData must outlive the
//    CacheManager built from it.
//  * to_kernel_object() returns a K_CacheManager<T> whose device API
//    (get / get_normalized) is declared in gc_cache_device.h; it is available
//    in code compiled by hipcc.
//  * T is the reference's voxel type (uchar1, ushort1, float1) or the plain
//    uint8_t, uint16_t, float.
#pragma once

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gc_cache_device.h"
#include "gcvol.h"

namespace tdns {

namespace math {
// tdns::math::Vector3ui (libMath/Vector.hpp), the three-component subset.
struct Vector3ui {
    uint32_t v[3] = {0, 0, 0};
    Vector3ui() = default;
    explicit Vector3ui(uint32_t s) : v{s, s, s} {}
    Vector3ui(uint32_t x, uint32_t y, uint32_t z) : v{x, y, z} {}
    uint32_t &operator[](size_t i) { return v[i]; }
    const uint32_t &operator[](size_t i) const { return v[i]; }
};
} // namespace math

namespace detail {
inline void check(int rc)
{
    if (rc != GC_OK) throw std::runtime_error(gc_last_error());
}
} // namespace detail

namespace data {

// Bkey (BrickKey.hpp, BrickKey.cpp:15-55): level, x, y, z as 16-bit fields.
typedef uint64_t Bkey;
inline Bkey get_key(uint32_t level, const math::Vector3ui &position)
{
    return gc_brick_key(level, position[0], position[1], position[2]);
}
inline void get_brick_level_position(Bkey key, uint32_t &level, math::Vector3ui &position)
{
    gc_brick_key_unpack(key, &level, position.v);
}
inline std::string get_brick_name(uint32_t level, const math::Vector3ui &position)
{
    return "L" + std::to_string(level) + "_" + std::to_string(position[0]) + "_" + std::to_string(position[1]) + "_" +
           std::to_string(position[2]);
}
inline std::string get_brick_name(Bkey key)
{
    uint32_t level = 0;
    math::Vector3ui p;
    get_brick_level_position(key, level, p);
    return get_brick_name(level, p);
}

struct VolumeConfiguration {
    std::string VolumeDirectory;
    std::string VolumeFileName;
    math::Vector3ui BrickSize;
    math::Vector3ui BigBrickSize;
    math::Vector3ui Covering;
    uint32_t EncodedBytes = 1;
    uint32_t Channels = 1;
    uint32_t NbLevels = 0;
    std::vector<math::Vector3ui> InitialVolumeSizes;
    std::vector<math::Vector3ui> RealVolumesSizes;
    std::vector<math::Vector3ui> NbBricks;
    std::vector<math::Vector3ui> NbBigBricks;
    std::vector<uint64_t> EmptyBricks;   // Bkey (BrickKey.cpp:15-25)
    std::vector<float> Histogram;
};

struct CacheConfiguration {
    std::vector<math::Vector3ui> CacheSize;   // [0]: brick slots per axis
    std::vector<math::Vector3ui> BlockSize;   // [0]: the brick size
    uint32_t DataCacheFlags = 0;              // 1 = NormalizedAccess
};

inline gc_volume_cfg to_gc(const VolumeConfiguration &v)
{
    gc_volume_cfg c{};
    for (int a = 0; a < 3; ++a) {
        c.brick_size[a] = v.BrickSize[a];
        c.big_brick_size[a] = v.BigBrickSize[a];
        c.covering[a] = v.Covering[a];
    }
    c.encoded_bytes = v.EncodedBytes;
    c.channels = v.Channels;
    c.nb_levels = v.NbLevels;
    if (v.NbLevels > GC_MAX_LEVELS || v.InitialVolumeSizes.size() < v.NbLevels || v.RealVolumesSizes.size() < v.NbLevels)
        throw std::runtime_error("VolumeConfiguration: level lists shorter than NbLevels");
    for (uint32_t l = 0; l < v.NbLevels; ++l)
        for (int a = 0; a < 3; ++a) {
            c.initial_sizes[l][a] = v.InitialVolumeSizes[l][a];
            c.real_sizes[l][a] = v.RealVolumesSizes[l][a];
        }
    return c;
}

inline VolumeConfiguration from_gc(const gc_volume_cfg &c)
{
    VolumeConfiguration v;
    v.BrickSize = math::Vector3ui(c.brick_size[0], c.brick_size[1], c.brick_size[2]);
    v.BigBrickSize = math::Vector3ui(c.big_brick_size[0], c.big_brick_size[1], c.big_brick_size[2]);
    v.Covering = math::Vector3ui(c.covering[0], c.covering[1], c.covering[2]);
    v.EncodedBytes = c.encoded_bytes;
    v.Channels = c.channels;
    v.NbLevels = c.nb_levels;
    for (uint32_t l = 0; l < c.nb_levels; ++l) {
        v.InitialVolumeSizes.emplace_back(c.initial_sizes[l][0], c.initial_sizes[l][1], c.initial_sizes[l][2]);
        v.RealVolumesSizes.emplace_back(c.real_sizes[l][0], c.real_sizes[l][1], c.real_sizes[l][2]);
        math::Vector3ui nb, nbb;
        for (int a = 0; a < 3; ++a) {
            nb[a] = c.real_sizes[l][a] / c.brick_size[a];
            nbb[a] = nb[a] / c.big_brick_size[a];
        }
        v.NbBricks.push_back(nb);
        v.NbBigBricks.push_back(nbb);
    }
    return v;
}

// load_volume_configuration (VolumeConfiguration.cpp:14-110).
inline VolumeConfiguration load_volume_configuration(const std::string &configurationFile)
{
    gc_volume_cfg c{};
    char path[4096];
    detail::check(gc_volume_cfg_load(configurationFile.c_str(), &c, path, sizeof(path)));
    VolumeConfiguration v = from_gc(c);
    const std::string p(path);
    const size_t slash = p.find_last_of("/\\");
    v.VolumeFileName = slash == std::string::npos ? p : p.substr(slash + 1);
    v.VolumeDirectory = slash == std::string::npos ? std::string() : p.substr(0, slash + 1);
    return v;
}

// MetaData (MetaData.hpp:16-103) plus the brick source behind it.
class MetaData {
public:
    MetaData() = default;
    MetaData(const MetaData &) = delete;
    MetaData &operator=(const MetaData &) = delete;
    ~MetaData() { reset(); }

    // Attach a brick source (borrowed unless take_ownership).
    void set_source(gc_source *src, bool take_ownership = false)
    {
        reset();
        _source = src;
        _owned = take_ownership;
    }
    gc_source *source() const { return _source; }

    // Read a raw volume (x-fastest, EncodedBytes per voxel, size
    // InitialVolumeSizes[0]), build its mip levels on `device` and serve
    // bricks from the in-memory pyramid.
    bool load_raw_volume(const VolumeConfiguration &vc, const std::string &path, int device = 0)
    {
        const gc_volume_cfg c = to_gc(vc);
        const uint64_t bytes = c.encoded_bytes;
        std::vector<std::vector<uint8_t>> levels(c.nb_levels);
        auto voxels = [&](uint32_t l) {
            return (uint64_t)c.initial_sizes[l][0] * c.initial_sizes[l][1] * c.initial_sizes[l][2];
        };
        levels[0].resize(voxels(0) * bytes);
        std::FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) return false;
        const size_t got = std::fread(levels[0].data(), 1, levels[0].size(), f);
        std::fclose(f);
        if (got != levels[0].size()) return false;
        detail::check(gc_set_device(device));
        void *d_in = nullptr, *d_out = nullptr;
        detail::check(gc_malloc(&d_in, levels[0].size()));
        try {
            detail::check(gc_memcpy_h2d(d_in, levels[0].data(), levels[0].size()));
            for (uint32_t l = 1; l < c.nb_levels; ++l) {
                for (int a = 0; a < 3; ++a)
                    if (c.initial_sizes[l][a] != c.initial_sizes[l - 1][a] / 2)
                        throw std::runtime_error("MetaData: level sizes do not follow the 2x2x2 reduction");
                levels[l].resize(voxels(l) * bytes);
                detail::check(gc_malloc(&d_out, levels[l].size() ? levels[l].size() : 1));
                detail::check(gc_mip_reduce(d_in, c.initial_sizes[l - 1], c.encoded_bytes, d_out, nullptr));
                detail::check(gc_memcpy_d2h(levels[l].data(), d_out, levels[l].size()));
                detail::check(gc_free(d_in));
                d_in = d_out;
                d_out = nullptr;
            }
        } catch (...) {
            gc_free(d_in);
            gc_free(d_out);
            throw;
        }
        detail::check(gc_free(d_in));
        std::vector<const void *> ptrs;
        for (auto &lv : levels) ptrs.push_back(lv.data());
        gc_source *src = nullptr;
        detail::check(gc_source_create_pyramid(&c, ptrs.data(), &src));
        reset();
        _levels = std::move(levels);
        _source = src;
        _owned = true;
        return true;
    }

    // MetaData::load (MetaData.cpp:106-213) plus the BricksManager over the
    // store: the store of the volume under `volumeDirectory` (default: the
    // configuration's VolumeDirectory + the volume file's stem, where
    // load_volume_configuration found bricks.xml), read through a host LRU
    // of cacheSize big bricks.
    bool load(const VolumeConfiguration &vc, std::string volumeDirectory = std::string(), size_t cacheSize = 32768)
    {
        if (volumeDirectory.empty()) {
            const std::string f = vc.VolumeFileName;
            const size_t dot = f.find_last_of('.');
            volumeDirectory = vc.VolumeDirectory + f.substr(0, dot) + "/";
        }
        const gc_volume_cfg c = to_gc(vc);
        uint32_t ne = 0, nh = 0;
        if (gc_store_read_meta(volumeDirectory.c_str(), c.brick_size[0], nullptr, &ne, nullptr, &nh) != GC_OK)
            return false;
        std::vector<uint64_t> empty(ne);
        std::vector<float> histo(nh);
        detail::check(gc_store_read_meta(volumeDirectory.c_str(), c.brick_size[0], empty.data(), &ne, histo.data(), &nh));
        gc_source *src = nullptr;
        detail::check(gc_source_create_store(&c, volumeDirectory.c_str(), (uint32_t)cacheSize, &src));
        reset();
        _source = src;
        _owned = true;
        _empty = std::move(empty);
        _histo = std::move(histo);
        return true;
    }

    std::vector<uint64_t> &get_empty_bricks() { return _empty; }
    const std::vector<uint64_t> &get_empty_bricks() const { return _empty; }
    std::vector<float> &get_histo() { return _histo; }
    const std::vector<float> &get_histo() const { return _histo; }

private:
    void reset()
    {
        if (_owned && _source) gc_source_destroy(_source);
        _source = nullptr;
        _owned = false;
        _levels.clear();
    }
    gc_source *_source = nullptr;
    bool _owned = false;
    std::vector<std::vector<uint8_t>> _levels;
    std::vector<uint64_t> _empty;
    std::vector<float> _histo;
};

} // namespace data

namespace gpucache {

enum VoxelStatus { Mapped = 0, Unmapped, Empty };   // K_CacheManager.hpp:24-29

// Voxel storage type of a CacheManager<T>.
template <typename T> struct voxel_of;
template <> struct voxel_of<uint8_t> { using type = uint8_t; };
template <> struct voxel_of<uint16_t> { using type = uint16_t; };
template <> struct voxel_of<float> { using type = float; };
#if defined(HIP_INCLUDE_HIP_AMD_DETAIL_HIP_VECTOR_TYPES_H)
template <> struct voxel_of<uchar1> { using type = uint8_t; };
template <> struct voxel_of<ushort1> { using type = uint16_t; };
template <> struct voxel_of<float1> { using type = float; };
#endif

// K_CacheManager<T> (K_CacheManager.hpp:34-133): passed by value to kernels.
template <typename T>
class K_CacheManager {
public:
    using voxel_type = typename voxel_of<T>::type;
    K_CacheManager() = default;
    explicit K_CacheManager(const gc_kcache &k) : _k(k) {}
    const gc_kcache &raw() const { return _k; }
#if defined(__HIPCC__)
    template <typename U>
    __device__ VoxelStatus get_normalized(uint32_t level, const float3 &position, U &output)
    {
        float v = 0.f;
        const gc_voxel_status s = gc_get_normalized<voxel_type>(_k, level, position, v);
        if (s != GC_VOXEL_EMPTY) output = static_cast<U>(v);
        return static_cast<VoxelStatus>(s);
    }
    __device__ VoxelStatus get(uint32_t level, const float3 &position, voxel_type &output)
    {
        return static_cast<VoxelStatus>(gc_get<voxel_type>(_k, level, position, output));
    }
#endif
private:
    gc_kcache _k{};
};

// CacheManager<T> (CacheManager.hpp:38-128).
template <typename T>
class CacheManager {
public:
    using voxel_type = typename voxel_of<T>::type;

    CacheManager(const data::VolumeConfiguration &volumeConfiguration, const data::MetaData &metaData,
                 const data::CacheConfiguration &cacheConfiguration, int32_t gpuID = 0)
    {
        if (volumeConfiguration.EncodedBytes != sizeof(voxel_type))
            throw std::runtime_error("CacheManager: NumberEncodedBytes does not match the voxel type");
        if (cacheConfiguration.CacheSize.empty() || cacheConfiguration.BlockSize.empty())
            throw std::runtime_error("CacheManager: CacheConfiguration needs CacheSize[0] and BlockSize[0]");
        if (cacheConfiguration.CacheSize.size() != cacheConfiguration.BlockSize.size())
            throw std::runtime_error("CacheManager: CacheSize and BlockSize need one entry per cache");
        if (cacheConfiguration.CacheSize.size() > 1 + GC_MAX_TABLES)
            throw std::runtime_error("CacheManager: at most GC_MAX_TABLES page-table caches");
        if (!metaData.source()) throw std::runtime_error("CacheManager: MetaData has no brick source");
        const gc_volume_cfg vc = data::to_gc(volumeConfiguration);
        gc_cache_cfg cc{};
        for (int a = 0; a < 3; ++a) {
            cc.cache_size[a] = cacheConfiguration.CacheSize[0][a];
            cc.block_size[a] = cacheConfiguration.BlockSize[0][a];
        }
        // CacheSize[1..] / BlockSize[1..]: the page-table caches between the
        // MRPD and the pool (CacheManager.hpp:151-236)
        cc.nb_tables = (uint32_t)cacheConfiguration.CacheSize.size() - 1;
        for (uint32_t k = 0; k < cc.nb_tables; ++k)
            for (int a = 0; a < 3; ++a) {
                cc.table_cache_size[k][a] = cacheConfiguration.CacheSize[k + 1][a];
                cc.table_block_size[k][a] = cacheConfiguration.BlockSize[k + 1][a];
            }
        cc.data_cache_flags = cacheConfiguration.DataCacheFlags;
        cc.max_requests = 50;   // _nbMaxRequests (CacheManager.hpp:263)
        cc.sync_requests = 0;   // requests served by the worker thread
        detail::check(gc_cache_create(&vc, &cc, metaData.source(), gpuID, &_cache));
    }
    ~CacheManager()
    {
        if (_cache) gc_cache_destroy(_cache);
    }
    CacheManager(const CacheManager &) = delete;
    CacheManager &operator=(const CacheManager &) = delete;

    // Kernel object for the current frame (valid until the next update()).
    K_CacheManager<T> to_kernel_object()
    {
        gc_cache_view v{};
        detail::check(gc_cache_view_get(_cache, &v));
        return K_CacheManager<T>(gc_kcache_from_view(&v));
    }

    // true when the last frame raised no brick request (CacheManager.hpp:319-354).
    bool update()
    {
        int done = 0;
        detail::check(gc_cache_update(_cache, &done));
        return done != 0;
    }

    // Pool occupancy in [0, 1] (CacheManager.hpp:358-383).
    void completude(std::vector<float> &vec)
    {
        float f = 0.f;
        detail::check(gc_cache_completude(_cache, &f));
        vec.assign(1, f);
    }

    gc_cache *handle() const { return _cache; }

private:
    gc_cache *_cache = nullptr;
};

} // namespace gpucache
} // namespace tdns
