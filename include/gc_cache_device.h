/* gc_cache_device.h -- device-side page-table API over a gc_cache_view, for
 * user HIP kernels (the reference's K_CacheManager<T>,
 * GcCore/include/GcCore/cuda/libGPUCache/K_CacheManager.hpp:34-461).
 *
 * gc_kcache is a plain C POD, identical for host and device compilers, that a
 * kernel receives by value, like K_CacheManager<T>; it carries the
 * reference's one-entry page cache per thread (K_CacheManager.hpp:297-303).
 * The device functions (hipcc only) are templated on the voxel type T:
 * uint8_t, uint16_t or float (NumberEncodedBytes 1, 2, 4).
 *
 *   gc_get_normalized<T>(c, level, p, out)   get_normalized (:267-361)
 *       p in [0, 1)^3 of the level's initial volume.  Mapped: stamps the
 *       slot's usage with the frame timestamp and returns the mean of the
 *       2x2x2 texels {e-1, e}^3 around the element position e
 *       (K_CacheDevice.hpp:191-201,316-334 + the linear texture filter),
 *       divided by the voxel type's maximum for integer voxels.  Unmapped:
 *       raises a brick request at the real-volume position (:365-396) and
 *       sets out = 0.  Empty: out is left unchanged.
 *   gc_get<T>(c, level, p, out)              get (:164-262)
 *       Mapped: the texel at e.  Unmapped: raises the request for the asked
 *       level, then tries the next coarser levels; Unmapped (out = 0) if no
 *       level is mapped.  Empty: out is left unchanged.
 *
 * With page-table caches (CacheSize[1..]) both walk from the MRPD entry
 * through the tables (:206-231,310-339), stamping each table block they read;
 * an Unmapped entry at any level of the walk counts as Unmapped.
 */
#ifndef GC_CACHE_DEVICE_H
#define GC_CACHE_DEVICE_H

#include "gcvol.h"

typedef enum gc_voxel_status {
    GC_VOXEL_MAPPED = 0,     /* VoxelStatus::Mapped   (K_CacheManager.hpp:24-29) */
    GC_VOXEL_UNMAPPED = 1,   /* VoxelStatus::Unmapped */
    GC_VOXEL_EMPTY = 2       /* VoxelStatus::Empty    */
} gc_voxel_status;

typedef struct gc_kcache {
    const uint32_t *mrpd;        /* uint4 entries (x, y, z, w), top_cells_dim, x-fastest */
    uint32_t *request;
    uint32_t *usage;
    const void *pool;            /* brick-major: slot s holds bs^3 voxels          */
    const gc_geometry *geom;     /* device copy                                    */
    uint32_t timestamp;
    uint32_t prev_coords[3];     /* one-entry page cache of the calling thread     */
    uint32_t prev_entry[4];
    uint32_t nb_tables;          /* page-table caches (gc_cache_view)              */
    const uint32_t *table_words[GC_MAX_TABLES];
    uint32_t *table_usage[GC_MAX_TABLES];
} gc_kcache;

/* Kernel object of a cache view (valid until the next update or destroy). */
static inline gc_kcache gc_kcache_from_view(const gc_cache_view *v)
{
    gc_kcache c;
    c.mrpd = (const uint32_t *)v->mrpd;
    c.request = v->request;
    c.usage = v->usage;
    c.pool = v->pool;
    c.geom = (const gc_geometry *)v->geometry;
    c.timestamp = v->timestamp;
    c.prev_coords[0] = c.prev_coords[1] = c.prev_coords[2] = 0xffffffffu;
    c.prev_entry[0] = c.prev_entry[1] = c.prev_entry[2] = c.prev_entry[3] = 0;
    c.nb_tables = v->nb_tables;
    for (int i = 0; i < GC_MAX_TABLES; ++i) {
        c.table_words[i] = v->table_words[i];
        c.table_usage[i] = v->table_usage[i];
    }
    return c;
}

#if defined(__cplusplus) && defined(__HIPCC__)
#include <hip/hip_runtime.h>

namespace gc_device {

/* initialOverReal, then realDivVirtual (K_CacheManager.hpp:276-290) */
__device__ inline void positions(const gc_geometry &g, uint32_t level, const float3 &p, float3 &real, float3 &virt)
{
    real = make_float3(p.x * g.ior[level][0], p.y * g.ior[level][1], p.z * g.ior[level][2]);
    virt = make_float3(real.x * g.rdv[level][0], real.y * g.rdv[level][1], real.z * g.rdv[level][2]);
}

/* K_MultiResolutionPageDirectory::get_coords (K_MultiResolutionPageDirectory.hpp:168-180) over the
 * brick grid: the request buffer's (raise_brick_request, K_CacheManager.hpp:365-396) */
__device__ inline uint3 page_coords(const gc_geometry &g, uint32_t level, const float3 &q)
{
    return make_uint3(g.level_coord[level] + (uint32_t)(q.x * (float)g.level_dim[level][0]),
                      (uint32_t)(q.y * (float)g.level_dim[level][1]),
                      (uint32_t)(q.z * (float)g.level_dim[level][2]));
}

/* request-buffer cell of brick-grid coordinates */
__device__ inline uint64_t cell(const gc_geometry &g, const uint3 &c)
{
    return c.x + (uint64_t)g.mrpd_dim[0] * (c.y + (uint64_t)g.mrpd_dim[1] * c.z);
}

/* MRPD coordinates (the top of the hierarchy: the brick grid without
 * page-table caches) */
__device__ inline uint3 top_coords(const gc_geometry &g, uint32_t level, const float3 &q)
{
    return make_uint3(g.top_coord[level] + (uint32_t)(q.x * (float)g.top_dim[level][0]),
                      (uint32_t)(q.y * (float)g.top_dim[level][1]),
                      (uint32_t)(q.z * (float)g.top_dim[level][2]));
}

__device__ inline uint4 entry_at(const gc_kcache &k, const uint3 &c)
{
    const gc_geometry &g = *k.geom;
    return reinterpret_cast<const uint4 *>(k.mrpd)[c.x + (uint64_t)g.top_cells_dim[0] *
                                                             (c.y + (uint64_t)g.top_cells_dim[1] * c.z)];
}

/* From a Mapped MRPD entry through the page-table caches
 * (K_CacheManager.hpp:206-231,310-339): each table read stamps its block's
 * usage (compute_element_position, K_CacheDevice.hpp:316-327) and yields the
 * next entry, (origin of the block or brick slot it maps, w); the walk ends
 * at the first entry that is not Mapped. */
__device__ inline uint4 walk_tables(const gc_kcache &k, uint4 e, uint32_t level, const float3 &virt)
{
    const gc_geometry &g = *k.geom;
    for (uint32_t i = 0; i < k.nb_tables && e.w == 1; ++i) {
        const gc_table_geometry &t = g.table[i];
        const uint32_t b = e.x / t.block[0] + t.blocks[0] * (e.y / t.block[1] + t.blocks[1] * (e.z / t.block[2]));
        k.table_usage[i][b] = k.timestamp;
        const uint32_t lx = (uint32_t)(virt.x * (float)t.virt[level][0]) % t.block[0];
        const uint32_t ly = (uint32_t)(virt.y * (float)t.virt[level][1]) % t.block[1];
        const uint32_t lz = (uint32_t)(virt.z * (float)t.virt[level][2]) % t.block[2];
        const uint32_t w = k.table_words[i][(uint64_t)b * (t.block[0] * t.block[1] * t.block[2]) + lx +
                                            t.block[0] * (ly + t.block[1] * lz)];
        const uint32_t c = w >> 2;
        uint32_t o[3] = {0, 0, 0};
        if ((w & 3u) == 1) {
            if (i + 1 < k.nb_tables) {
                const gc_table_geometry &n = g.table[i + 1];
                o[0] = (c % n.blocks[0]) * n.block[0];
                o[1] = ((c / n.blocks[0]) % n.blocks[1]) * n.block[1];
                o[2] = (c / (n.blocks[0] * n.blocks[1])) * n.block[2];
            } else {
                o[0] = (c % g.cache_size[0]) * g.brick_size;
                o[1] = ((c / g.cache_size[0]) % g.cache_size[1]) * g.brick_size;
                o[2] = (c / (g.cache_size[0] * g.cache_size[1])) * g.brick_size;
            }
        }
        e = make_uint4(o[0], o[1], o[2], w & 3u);
    }
    return e;
}

/* raise_brick_request (K_CacheManager.hpp:365-396) */
__device__ inline void raise_request(const gc_kcache &k, uint32_t level, const float3 &real)
{
    k.request[cell(*k.geom, page_coords(*k.geom, level, real))] = k.timestamp;
}

/* Usage stamp and element position in the pool's 3-D voxel space
 * (K_CacheDevice::compute_element_position, K_CacheDevice.hpp:316-334). */
__device__ inline uint3 element(const gc_kcache &k, const uint4 &e, uint32_t level, const float3 &virt)
{
    const gc_geometry &g = *k.geom;
    const uint32_t bs = g.brick_size, cov = g.covering, inner = g.inner;
    k.usage[e.x / bs + g.cache_size[0] * (e.y / bs + g.cache_size[1] * (e.z / bs))] = k.timestamp;
    return make_uint3(cov + e.x + (uint32_t)(virt.x * (float)g.virt_vox[level][0]) % inner,
                      cov + e.y + (uint32_t)(virt.y * (float)g.virt_vox[level][1]) % inner,
                      cov + e.z + (uint32_t)(virt.z * (float)g.virt_vox[level][2]) % inner);
}

/* Texel at 3-D pool coordinates, x and y clamped, z wrapped (the pool
 * texture's address modes, Texture3dDevice.hpp:104-122). */
template <typename T>
__device__ inline T texel(const gc_kcache &k, int x, int y, int z)
{
    const gc_geometry &g = *k.geom;
    const uint32_t bs = g.brick_size;
    const int PX = (int)(g.cache_size[0] * bs), PY = (int)(g.cache_size[1] * bs), PZ = (int)(g.cache_size[2] * bs);
    x = x < 0 ? 0 : (x >= PX ? PX - 1 : x);
    y = y < 0 ? 0 : (y >= PY ? PY - 1 : y);
    z = z < 0 ? PZ - 1 : (z >= PZ ? z - PZ : z);
    const uint32_t slot = (uint32_t)x / bs + g.cache_size[0] * ((uint32_t)y / bs + g.cache_size[1] * ((uint32_t)z / bs));
    return static_cast<const T *>(k.pool)[(uint64_t)slot * bs * bs * bs + (uint32_t)x % bs +
                                          (uint64_t)bs * ((uint32_t)y % bs + (uint64_t)bs * ((uint32_t)z % bs))];
}

/* Linear filter at an element position: the mean of {e-1, e}^3, normalised
 * by the integer type's maximum. */
template <typename T>
__device__ inline float box(const gc_kcache &k, const uint3 &el)
{
    float s = 0.f;
    for (int kk = 0; kk < 2; ++kk)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i) s += (float)texel<T>(k, (int)el.x - 1 + i, (int)el.y - 1 + j, (int)el.z - 1 + kk);
    if (sizeof(T) == 4) return s / 8.0f;
    return s * (1.0f / (8.0f * (sizeof(T) == 1 ? 255.f : 65535.f)));
}

} // namespace gc_device

template <typename T>
__device__ inline gc_voxel_status gc_get_normalized(gc_kcache &k, uint32_t level, const float3 &p, float &out)
{
    using namespace gc_device;
    float3 real, virt;
    positions(*k.geom, level, p, real, virt);
    const uint3 c = top_coords(*k.geom, level, virt);
    if (c.x != k.prev_coords[0] || c.y != k.prev_coords[1] || c.z != k.prev_coords[2]) {
        const uint4 e = entry_at(k, c);
        k.prev_entry[0] = e.x; k.prev_entry[1] = e.y; k.prev_entry[2] = e.z; k.prev_entry[3] = e.w;
        k.prev_coords[0] = c.x; k.prev_coords[1] = c.y; k.prev_coords[2] = c.z;
    }
    /* an Unmapped table entry is Unmapped (the reference, with a coarser
     * level pending, repeats the same lookup for ever there,
     * K_CacheManager.hpp:318-330) */
    const uint4 e = walk_tables(k, make_uint4(k.prev_entry[0], k.prev_entry[1], k.prev_entry[2], k.prev_entry[3]),
                                level, virt);
    if (e.w == 1) {
        out = box<T>(k, element(k, e, level, virt));
        return GC_VOXEL_MAPPED;
    }
    if (e.w == 0) {
        raise_request(k, level, real);
        out = 0.f;
        return GC_VOXEL_UNMAPPED;
    }
    return GC_VOXEL_EMPTY;
}

template <typename T>
__device__ inline gc_voxel_status gc_get(const gc_kcache &k, uint32_t level, const float3 &p, T &out)
{
    using namespace gc_device;
    const uint32_t asked = level, last = k.geom->nb_levels - 1;
    for (;;) {
        float3 real, virt;
        positions(*k.geom, level, p, real, virt);
        /* an Unmapped entry at any level of the walk falls back to the
         * coarser level (the reference does so only for the MRPD's; after a
         * table's it repeats the lookup for ever, K_CacheManager.hpp:212-222) */
        const uint4 e = walk_tables(k, entry_at(k, top_coords(*k.geom, level, virt)), level, virt);
        if (e.w == 1) {
            const uint3 el = element(k, e, level, virt);
            out = texel<T>(k, (int)el.x, (int)el.y, (int)el.z);
            return GC_VOXEL_MAPPED;
        }
        if (e.w == 2) return GC_VOXEL_EMPTY;
        if (level == asked) raise_request(k, level, real);
        if (level < last) {
            ++level;
            continue;
        }
        out = T(0);
        return GC_VOXEL_UNMAPPED;
    }
}
#endif /* __cplusplus && __HIPCC__ */

#endif /* GC_CACHE_DEVICE_H */
