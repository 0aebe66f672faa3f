// kernels_cache.hip -- cache bookkeeping and preprocessing kernels for gfx950.
//
//  * LRU update: one stable 1-bit radix pass keyed on "used this frame"
//    (create_usage_mask + the two thrust::copy_if of CacheDevice::update,
//    CacheManagerKernels.cu:8-29, CacheDevice.hpp:202-219), built from wave
//    ballots and popcount prefix sums.
//  * Miss compaction: stable compaction of request cells == timestamp over
//    [0, cells-1) with the same ballot/scan primitives, decoded like
//    coord_1D_to_3D (CacheManager.hpp:410-474, CacheManagerKernels.cu:32-72).
//  * Eviction and page-table update (CacheManager.hpp:503-632 and the
//    kernels of CacheManagerKernels.hpp:59-357), one workgroup per brick.
//  * 2x2x2 mip reduction (Mipmapper.cu:30-105) and brick cutting with
//    covering (Bricker_v2.cpp:79-252).
#include "gc_internal.hpp"

namespace gc {

namespace {

constexpr int BLK = 256;             // 4 waves of 64
constexpr int ROUNDS = 4;
constexpr int TILE = BLK * ROUNDS;   // entries per workgroup, 16 ballot words
constexpr int WORDS = TILE / 64;

__device__ inline uint32_t popc64(uint64_t m) { return (uint32_t)__popcll(m); }

// ----------------------------------------------------------------- LRU pass
__global__ __launch_bounds__(BLK) void k_lru_flags(const uint32_t *__restrict__ order,
                                                   const uint32_t *__restrict__ usage, uint32_t n,
                                                   uint32_t ts, uint64_t *__restrict__ masks,
                                                   uint32_t *__restrict__ counts)
{
    __shared__ uint32_t wc[WORDS];
    const uint32_t base = blockIdx.x * TILE;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const uint32_t i = base + r * BLK + threadIdx.x;
        bool used = false;
        if (i < n) used = usage[order[i]] == ts;
        const uint64_t m = __ballot(used);
        const uint32_t word = (base + r * BLK) / 64 + w;
        if (lane == 0) {
            if ((uint64_t)word * 64 < n) masks[word] = m;
            wc[r * 4 + w] = popc64(m);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int k = 0; k < WORDS; ++k) s += wc[k];
        counts[blockIdx.x] = s;
    }
}

// Exclusive scan of per-workgroup counts in place, total in total[0].
__global__ __launch_bounds__(1024) void k_scan(uint32_t *counts, uint32_t nb, uint32_t *total)
{
    __shared__ uint32_t sh[1024];
    const uint32_t t = threadIdx.x;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t i = base + t;
        const uint32_t v = i < nb ? counts[i] : 0;
        sh[t] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 1024; off <<= 1) {
            const uint32_t x = t >= off ? sh[t - off] : 0;
            __syncthreads();
            sh[t] += x;
            __syncthreads();
        }
        const uint32_t incl = sh[t];
        const uint32_t tot = sh[1023];
        if (i < nb) counts[i] = carry + incl - v;
        carry += tot;
        __syncthreads();
    }
    if (t == 0) total[0] = carry;
}

// Prefix of set bits of the tile's ballot words before word k, in LDS.
__device__ inline void tile_word_prefix(const uint64_t *masks, uint32_t base, uint64_t n,
                                        uint32_t *wpre)
{
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int k = 0; k < WORDS; ++k) {
            wpre[k] = s;
            const uint64_t word = base / 64 + k;
            if (word * 64 < n) s += popc64(masks[word]);
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(BLK) void k_lru_scatter(const uint32_t *__restrict__ in,
                                                     uint32_t *__restrict__ out, uint32_t n,
                                                     const uint64_t *__restrict__ masks,
                                                     const uint32_t *__restrict__ offsets,
                                                     const uint32_t *__restrict__ total)
{
    __shared__ uint32_t wpre[WORDS];
    const uint32_t base = blockIdx.x * TILE;
    tile_word_prefix(masks, base, n, wpre);
    const uint32_t off = offsets[blockIdx.x];
    const uint32_t used_total = total[0];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const uint32_t i = base + r * BLK + threadIdx.x;
        if (i >= n) continue;
        const uint64_t m = masks[i / 64];
        const uint32_t lane = i & 63;
        const uint32_t below = popc64(m & ((1ull << lane) - 1ull));
        const uint32_t before = off + wpre[(i - base) / 64] + below;   // used entries before i
        const bool used = (m >> lane) & 1ull;
        const uint32_t dst = used ? before : used_total + (i - before);
        out[dst] = in[i];
    }
}

// ------------------------------------------------------------ miss compaction
__global__ __launch_bounds__(BLK) void k_miss_flags(const uint32_t *__restrict__ req, uint32_t n,
                                                    uint32_t ts, uint64_t *__restrict__ masks,
                                                    uint32_t *__restrict__ counts)
{
    __shared__ uint32_t wc[WORDS];
    const uint32_t base = blockIdx.x * TILE;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const uint32_t i = base + r * BLK + threadIdx.x;
        const bool hit = i < n && req[i] == ts;
        const uint64_t m = __ballot(hit);
        const uint32_t word = (base + r * BLK) / 64 + w;
        if (lane == 0) {
            if ((uint64_t)word * 64 < n) masks[word] = m;
            wc[r * 4 + w] = popc64(m);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int k = 0; k < WORDS; ++k) s += wc[k];
        counts[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(BLK) void k_miss_scatter(uint32_t n, const uint64_t *__restrict__ masks,
                                                      const uint32_t *__restrict__ offsets,
                                                      const gc_geometry *__restrict__ g,
                                                      uint32_t cap, uint32_t *__restrict__ out4)
{
    __shared__ uint32_t wpre[WORDS];
    const uint32_t off = offsets[blockIdx.x];
    if (off >= cap) return;   // uniform per workgroup: every hit here ranks >= cap
    const uint32_t base = blockIdx.x * TILE;
    tile_word_prefix(masks, base, n, wpre);
    const uint32_t SX = g->mrpd_dim[0];
    const uint32_t SY = g->level_dim[0][1];
    for (int r = 0; r < ROUNDS; ++r) {
        const uint32_t i = base + r * BLK + threadIdx.x;
        if (i >= n) continue;
        const uint64_t m = masks[i / 64];
        const uint32_t lane = i & 63;
        if (!((m >> lane) & 1ull)) continue;
        const uint32_t rank = off + wpre[(i - base) / 64] + popc64(m & ((1ull << lane) - 1ull));
        if (rank >= cap) continue;
        // coord_1D_to_3D (CacheManagerKernels.cu:41-71)
        uint64_t idx = i;
        const uint32_t z = (uint32_t)(idx / ((uint64_t)SX * SY));
        idx -= (uint64_t)z * SX * SY;
        const uint32_t y = (uint32_t)(idx / SX);
        idx -= (uint64_t)y * SX;
        const uint32_t x = (uint32_t)idx;
        uint32_t sum = 0, level;
        for (level = 0; level < g->nb_levels; ++level) {
            sum += g->level_dim[level][0];
            if (x < sum) break;
        }
        uint4 o;
        o.x = x - (sum - g->level_dim[level][0]);
        o.y = y;
        o.z = z;
        o.w = level;
        reinterpret_cast<uint4 *>(out4)[rank] = o;
    }
}

// ------------------------------------------------------------------ loading
__global__ void k_deref(uint4 *__restrict__ mrpd, uint32_t *__restrict__ cells, const uint32_t *__restrict__ order,
                        uint32_t N, uint32_t nne, const SlotMeta *__restrict__ meta,
                        const gc_geometry *__restrict__ g)
{
    // dereference_old_bricks (CacheManagerKernels.hpp:115-145), bound j < nne.
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nne) return;
    const SlotMeta m = meta[order[N - nne + j]];
    if (m.vol[0] < 0.f) return;   // never-filled slot
    const uint64_t idx = mrpd_index(*g, m.level, m.vol);
    mrpd[idx] = make_uint4(0, 0, 0, 0);
    cells[idx] = 0u;
}

template <int BYTES>
__global__ __launch_bounds__(BLK) void k_assign_copy(const uint32_t *__restrict__ order, uint32_t N,
                                                     uint32_t nne, const uint32_t *__restrict__ asked4,
                                                     const uint32_t *__restrict__ bb_index,
                                                     const uint32_t *__restrict__ bb_coord4,
                                                     SlotMeta *__restrict__ meta, uint8_t *__restrict__ pool,
                                                     const uint8_t *__restrict__ staging,
                                                     const gc_geometry *__restrict__ g, uint3 BB)
{
    // compute_lru_volume_position + copy_lru_position + write_brick
    // (CacheManagerKernels.cu:90-114, CacheManagerKernels.hpp:59-107).
    const uint32_t k = blockIdx.x;
    const uint32_t s = order[N - nne + k];
    const uint32_t *a = asked4 + 4 * k;
    if (threadIdx.x == 0) {
        SlotMeta m;
        m.level = a[3];
        volume_position(*g, a[3], a, m.vol);
        meta[s] = m;
    }
    const uint32_t bs = g->brick_size;
    const uint64_t bvox = (uint64_t)bs * bs * bs;
    uint8_t *dst = pool + (uint64_t)s * bvox * BYTES;
    const uint32_t bbi = bb_index[k];
    if (BB.x == 1 && BB.y == 1 && BB.z == 1) {
        const uint8_t *src = staging + (uint64_t)bbi * bvox * BYTES;
        const uint64_t bytes = bvox * BYTES;
        if ((bytes & 15) == 0) {
            const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
            uint4 *d4 = reinterpret_cast<uint4 *>(dst);
            for (uint64_t i = threadIdx.x; i < bytes / 16; i += BLK) d4[i] = s4[i];
        } else {
            for (uint64_t i = threadIdx.x; i < bytes; i += BLK) dst[i] = src[i];
        }
        return;
    }
    const uint32_t *bbc = bb_coord4 + 4 * bbi;
    const uint64_t BVX = (uint64_t)BB.x * bs, BVY = (uint64_t)BB.y * bs, BVZ = (uint64_t)BB.z * bs;
    const uint64_t ox = a[0] - bbc[0] * BB.x, oy = a[1] - bbc[1] * BB.y, oz = a[2] - bbc[2] * BB.z;
    const uint64_t base = (uint64_t)bbi * BVX * BVY * BVZ + ((oz * BVY + oy) * BVX + ox) * bs;
    for (uint64_t v = threadIdx.x; v < bvox; v += BLK) {
        const uint64_t x = v % bs, y = (v / bs) % bs, z = v / ((uint64_t)bs * bs);
        const uint64_t src = base + (z * BVY + y) * BVX + x;
#pragma unroll
        for (int b = 0; b < BYTES; ++b) dst[v * BYTES + b] = staging[src * BYTES + b];
    }
}

template <int BYTES>
__global__ __launch_bounds__(BLK) void k_assign_cut(const uint32_t *__restrict__ order, uint32_t N,
                                                    uint32_t n, const uint32_t *__restrict__ asked4,
                                                    SlotMeta *__restrict__ meta, uint8_t *__restrict__ pool,
                                                    const void *const *__restrict__ levels,
                                                    const uint32_t *__restrict__ level_dims,
                                                    const gc_geometry *__restrict__ g)
{
    const uint32_t k = blockIdx.x;
    const uint32_t s = order[N - n + k];
    const uint32_t *a = asked4 + 4 * k;
    const uint32_t level = a[3];
    if (threadIdx.x == 0) {
        SlotMeta m;
        m.level = level;
        volume_position(*g, level, a, m.vol);
        meta[s] = m;
    }
    const uint32_t bs = g->brick_size, cov = g->covering, inner = g->inner;
    const uint64_t bvox = (uint64_t)bs * bs * bs;
    uint8_t *dst = pool + (uint64_t)s * bvox * BYTES;
    const uint8_t *src = static_cast<const uint8_t *>(levels[level]);
    const uint32_t X = level_dims[3 * level], Y = level_dims[3 * level + 1], Z = level_dims[3 * level + 2];
    const int64_t ox = (int64_t)a[0] * inner - cov, oy = (int64_t)a[1] * inner - cov,
                  oz = (int64_t)a[2] * inner - cov;
    for (uint64_t v = threadIdx.x; v < bvox; v += BLK) {
        const int64_t gx = ox + (int64_t)(v % bs), gy = oy + (int64_t)((v / bs) % bs),
                      gz = oz + (int64_t)(v / ((uint64_t)bs * bs));
        const bool in = gx >= 0 && gy >= 0 && gz >= 0 && gx < X && gy < Y && gz < Z;
        const uint64_t si = in ? (uint64_t)gx + (uint64_t)X * ((uint64_t)gy + (uint64_t)Y * gz) : 0;
#pragma unroll
        for (int b = 0; b < BYTES; ++b) dst[v * BYTES + b] = in ? src[si * BYTES + b] : 0;
    }
}

__device__ inline void insert_one(uint4 *mrpd, uint32_t *cells, const uint32_t *order, uint32_t N, uint32_t nne,
                                  const uint32_t *a, uint32_t k, const gc_geometry &g)
{
    // reference_new_bricks / upward for the MRPD-only hierarchy
    // (CacheManagerKernels.hpp:236-251,336-353): insert if Unmapped.
    float vol[3];
    volume_position(g, a[3], a, vol);
    const uint64_t idx = mrpd_index(g, a[3], vol);
    if (mrpd[idx].w >= 1) return;
    uint4 e;
    uint32_t c;
    if (k < nne) {
        const uint32_t slot = order[N - nne + k];
        uint32_t o[3];
        slot_origin(g, slot, o);
        e = make_uint4(o[0], o[1], o[2], 1);
        c = cell_word(slot, 1);
    } else {
        e = make_uint4(0, 0, 0, 2);   // zero-initialised draft position (CacheManager.hpp:546)
        c = cell_word(0, 2);
    }
    mrpd[idx] = e;
    cells[idx] = c;
}

__global__ void k_insert(uint4 *__restrict__ mrpd, uint32_t *__restrict__ cells, const uint32_t *__restrict__ order,
                         uint32_t N, uint32_t n, uint32_t nne, const uint32_t *__restrict__ asked4,
                         const gc_geometry *__restrict__ g)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    insert_one(mrpd, cells, order, N, nne, asked4 + 4 * k, k, *g);
}

__global__ void k_insert_serial(uint4 *mrpd, uint32_t *cells, const uint32_t *order, uint32_t N, uint32_t n,
                                uint32_t nne, const uint32_t *asked4, const gc_geometry *g)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (uint32_t k = 0; k < n; ++k) insert_one(mrpd, cells, order, N, nne, asked4 + 4 * k, k, *g);
}

__global__ __launch_bounds__(BLK) void k_occupancy(const SlotMeta *__restrict__ meta, uint32_t n,
                                                   uint32_t *__restrict__ count)
{
    // cache_use (CacheManagerKernels.cu:117-125), one atomic per wave.
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool filled = i < n && meta[i].vol[0] != -1.f;
    const uint64_t m = __ballot(filled);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, popc64(m));
}

// --------------------------------------------------------------- preprocess
template <typename T>
__global__ __launch_bounds__(BLK) void k_mip2(const T *__restrict__ in, uint32_t X, uint32_t Y,
                                              T *__restrict__ out, uint32_t OX, uint32_t OY,
                                              uint64_t total)
{
    // down_sampling_inner with a 2x2x2 group, summed z, y, x in that order
    // (Mipmapper.cu:68-103).
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += stride) {
        const uint64_t ox = o % OX, oy = (o / OX) % OY, oz = o / ((uint64_t)OX * OY);
        const uint64_t p = 2 * ox + (uint64_t)X * (2 * oy + (uint64_t)Y * 2 * oz);
        const uint64_t sx = 1, sy = X, sz = (uint64_t)X * Y;
        if constexpr (sizeof(T) == 4) {
            double s = 0.0;
            s += (double)in[p]; s += (double)in[p + sx];
            s += (double)in[p + sy]; s += (double)in[p + sy + sx];
            s += (double)in[p + sz]; s += (double)in[p + sz + sx];
            s += (double)in[p + sz + sy]; s += (double)in[p + sz + sy + sx];
            out[o] = (T)(s / 8.0);
        } else {
            const uint32_t s = (uint32_t)in[p] + in[p + sx] + in[p + sy] + in[p + sy + sx] +
                               in[p + sz] + in[p + sz + sx] + in[p + sz + sy] + in[p + sz + sy + sx];
            out[o] = (T)(s >> 3);
        }
    }
}

// One brick of level l built from level l-1 bricks (mip on miss): voxel g of
// level l (global, zero outside the level) = floor(sum of level-(l-1) voxels
// 2g + {0,1}^3 / 8) with k_mip2's types and order; each level-(l-1) voxel p
// is read from the inner region of the child brick p / inner that holds it.
template <typename T>
__global__ __launch_bounds__(BLK) void k_mip_brick(const MipBrickArgs a, const T *__restrict__ children,
                                                   T *__restrict__ out)
{
    const uint32_t bs = a.bs, n = bs * bs * bs;
    const uint64_t cvox = (uint64_t)bs * bs * bs;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int64_t g[3] = {a.g0[0] + (int64_t)(i % bs), a.g0[1] + (int64_t)((i / bs) % bs),
                              a.g0[2] + (int64_t)(i / (bs * bs))};
        bool in = true;
        for (int k = 0; k < 3; ++k) in = in && g[k] >= 0 && g[k] < (int64_t)a.dims[k];
        if (!in) {
            out[i] = T(0);
            continue;
        }
        double ds = 0.0;
        uint32_t is = 0;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const uint32_t p[3] = {(uint32_t)(2 * g[0] + dx), (uint32_t)(2 * g[1] + dy), (uint32_t)(2 * g[2] + dz)};
                    uint32_t c[3], l[3];
                    for (int k = 0; k < 3; ++k) {
                        c[k] = p[k] / a.inner - a.c0[k];
                        l[k] = p[k] % a.inner + a.cov;
                    }
                    const uint32_t ci = c[0] + a.nc[0] * (c[1] + a.nc[1] * c[2]);
                    const T v = children[ci * cvox + l[0] + (uint64_t)bs * (l[1] + (uint64_t)bs * l[2])];
                    if constexpr (sizeof(T) == 4) ds += (double)v;
                    else is += (uint32_t)v;
                }
        if constexpr (sizeof(T) == 4) out[i] = (T)(ds / 8.0);
        else out[i] = (T)(is >> 3);
    }
}

template <int BYTES>
__global__ __launch_bounds__(BLK) void k_brick_cut(const uint8_t *__restrict__ level, uint32_t X,
                                                   uint32_t Y, uint32_t Z, uint32_t bs, uint32_t cov,
                                                   const uint32_t *__restrict__ bricks,
                                                   uint8_t *__restrict__ out)
{
    const uint32_t k = blockIdx.x;
    const uint32_t inner = bs - 2 * cov;
    const uint64_t bvox = (uint64_t)bs * bs * bs;
    const int64_t ox = (int64_t)bricks[3 * k] * inner - cov, oy = (int64_t)bricks[3 * k + 1] * inner - cov,
                  oz = (int64_t)bricks[3 * k + 2] * inner - cov;
    uint8_t *dst = out + (uint64_t)k * bvox * BYTES;
    for (uint64_t v = threadIdx.x; v < bvox; v += BLK) {
        const int64_t gx = ox + (int64_t)(v % bs), gy = oy + (int64_t)((v / bs) % bs),
                      gz = oz + (int64_t)(v / ((uint64_t)bs * bs));
        const bool in = gx >= 0 && gy >= 0 && gz >= 0 && gx < X && gy < Y && gz < Z;
        const uint64_t si = in ? (uint64_t)gx + (uint64_t)X * ((uint64_t)gy + (uint64_t)Y * gz) : 0;
#pragma unroll
        for (int b = 0; b < BYTES; ++b) dst[v * BYTES + b] = in ? level[si * BYTES + b] : 0;
    }
}

// One workgroup per key: the brick of a resident level with the covering
// rule (k_brick_cut's), and the EmptyThreshold flag.
template <int BYTES>
__global__ __launch_bounds__(BLK) void k_cut_keys(const CutKeys a)
{
    const uint32_t u = blockIdx.x;
    const uint32_t *key = a.keys + 4ull * u;
    const uint32_t level = key[3];
    if (level < a.first_level) return;   // the whole workgroup: another writer's brick
    const uint32_t bs = a.bs, inner = bs - 2 * a.cov;
    const uint32_t X = a.dims[level][0], Y = a.dims[level][1], Z = a.dims[level][2];
    const uint8_t *src = a.levels[level];
    const uint64_t bvox = (uint64_t)bs * bs * bs;
    const int64_t ox = (int64_t)key[0] * inner - a.cov, oy = (int64_t)key[1] * inner - a.cov,
                  oz = (int64_t)key[2] * inner - a.cov;
    uint8_t *dst = a.out + (uint64_t)u * bvox * BYTES;
    int above = 0;
    for (uint64_t v = threadIdx.x; v < bvox; v += BLK) {
        const int64_t gx = ox + (int64_t)(v % bs), gy = oy + (int64_t)((v / bs) % bs),
                      gz = oz + (int64_t)(v / ((uint64_t)bs * bs));
        const bool in = gx >= 0 && gy >= 0 && gz >= 0 && gx < X && gy < Y && gz < Z;
        const uint64_t si = in ? (uint64_t)gx + (uint64_t)X * ((uint64_t)gy + (uint64_t)Y * gz) : 0;
        float val;
        if constexpr (BYTES == 1) {
            const uint8_t t = in ? src[si] : 0;
            dst[v] = t;
            val = (float)t;
        } else if constexpr (BYTES == 2) {
            const uint16_t t = in ? reinterpret_cast<const uint16_t *>(src)[si] : 0;
            reinterpret_cast<uint16_t *>(dst)[v] = t;
            val = (float)t;
        } else {
            const float t = in ? reinterpret_cast<const float *>(src)[si] : 0.f;
            reinterpret_cast<float *>(dst)[v] = t;
            val = t;
        }
        above |= !(val < a.thr);
    }
    above = __syncthreads_or(above);
    if (threadIdx.x == 0 && a.below) a.below[u] = above ? 0u : 1u;
}

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

} // namespace

// ---------------------------------------------------------------- launchers
hipError_t launch_lru_update(const uint32_t *order_in, uint32_t *order_out, const uint32_t *usage,
                             uint32_t n, uint32_t ts, const LruScratch &s, hipStream_t stream)
{
    const uint32_t nb = blocks_for(n, TILE);
    if (nb == 0) return hipSuccess;
    k_lru_flags<<<nb, BLK, 0, stream>>>(order_in, usage, n, ts, s.masks, s.counts);
    k_scan<<<1, 1024, 0, stream>>>(s.counts, nb, s.total);
    k_lru_scatter<<<nb, BLK, 0, stream>>>(order_in, order_out, n, s.masks, s.counts, s.total);
    return hipGetLastError();
}

hipError_t launch_miss_compact(const uint32_t *req, uint64_t cells, uint32_t ts,
                               const gc_geometry *d_geom, uint32_t cap, uint32_t *out4,
                               const MissScratch &s, hipStream_t stream)
{
    // thrust::copy_if over counting_iterator [0, size-1): the last cell is
    // excluded (CacheManager.hpp:415-422).
    const uint32_t n = cells > 0 ? (uint32_t)(cells - 1) : 0;
    const uint32_t nb = blocks_for(n, TILE);
    if (nb == 0) return hipMemsetAsync(s.total, 0, sizeof(uint32_t), stream);
    k_miss_flags<<<nb, BLK, 0, stream>>>(req, n, ts, s.masks, s.counts);
    k_scan<<<1, 1024, 0, stream>>>(s.counts, nb, s.total);
    if (cap > 0) k_miss_scatter<<<nb, BLK, 0, stream>>>(n, s.masks, s.counts, d_geom, cap, out4);
    return hipGetLastError();
}

hipError_t launch_deref(uint4 *mrpd, uint32_t *cells, const uint32_t *order, uint32_t nslots, uint32_t nne,
                        const SlotMeta *meta, const gc_geometry *d_geom, hipStream_t stream)
{
    if (nne == 0) return hipSuccess;
    k_deref<<<blocks_for(nne, BLK), BLK, 0, stream>>>(mrpd, cells, order, nslots, nne, meta, d_geom);
    return hipGetLastError();
}

hipError_t launch_assign_copy(const uint32_t *order, uint32_t nslots, uint32_t nne,
                              const uint32_t *asked4, const uint32_t *bb_index,
                              const uint32_t *bb_coord4, SlotMeta *meta, void *pool,
                              const void *staging, const gc_geometry *d_geom,
                              const uint32_t big_brick[3], uint32_t bytes, hipStream_t stream)
{
    if (nne == 0) return hipSuccess;
    const uint3 BB = make_uint3(big_brick[0], big_brick[1], big_brick[2]);
    uint8_t *p = static_cast<uint8_t *>(pool);
    const uint8_t *st = static_cast<const uint8_t *>(staging);
    switch (bytes) {
    case 1: k_assign_copy<1><<<nne, BLK, 0, stream>>>(order, nslots, nne, asked4, bb_index, bb_coord4, meta, p, st, d_geom, BB); break;
    case 2: k_assign_copy<2><<<nne, BLK, 0, stream>>>(order, nslots, nne, asked4, bb_index, bb_coord4, meta, p, st, d_geom, BB); break;
    case 4: k_assign_copy<4><<<nne, BLK, 0, stream>>>(order, nslots, nne, asked4, bb_index, bb_coord4, meta, p, st, d_geom, BB); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_assign_cut(const uint32_t *order, uint32_t nslots, uint32_t n,
                             const uint32_t *asked4, SlotMeta *meta, void *pool,
                             const void *const *d_levels, const uint32_t *d_level_dims,
                             const gc_geometry *d_geom, uint32_t bytes, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    uint8_t *p = static_cast<uint8_t *>(pool);
    switch (bytes) {
    case 1: k_assign_cut<1><<<n, BLK, 0, stream>>>(order, nslots, n, asked4, meta, p, d_levels, d_level_dims, d_geom); break;
    case 2: k_assign_cut<2><<<n, BLK, 0, stream>>>(order, nslots, n, asked4, meta, p, d_levels, d_level_dims, d_geom); break;
    case 4: k_assign_cut<4><<<n, BLK, 0, stream>>>(order, nslots, n, asked4, meta, p, d_levels, d_level_dims, d_geom); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_insert(uint4 *mrpd, uint32_t *cells, const uint32_t *order, uint32_t nslots, uint32_t n,
                         uint32_t nne, const uint32_t *asked4, const gc_geometry *d_geom,
                         int serial, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    if (serial) k_insert_serial<<<1, 64, 0, stream>>>(mrpd, cells, order, nslots, n, nne, asked4, d_geom);
    else k_insert<<<blocks_for(n, BLK), BLK, 0, stream>>>(mrpd, cells, order, nslots, n, nne, asked4, d_geom);
    return hipGetLastError();
}

// ----------------------------------------------------- page-table caches
// dereference_old_bricks (CacheManagerKernels.hpp:111-145) through the table
// caches: the walk to a victim slot's entry reads the MRPD and tables 0 ..
// T-2 (stamping their blocks, the reference's get calls); the entry in the
// last table is cleared if it still names the slot.  The reference clears
// the entry the walk reaches whatever it holds -- after the slot's table
// block was evicted, another brick's entry, or, from an Unmapped entry,
// one at the surface origin; here the walk stops at an entry that is not
// Mapped and the last entry is cleared only when it names the slot.  Victims
// are independent: an entry names one slot, so at most one thread clears it.
__global__ void k_deref_tables(uint32_t *__restrict__ cells, TableState tab, const uint32_t *__restrict__ order,
                               uint32_t N, uint32_t nne, const SlotMeta *__restrict__ meta,
                               const gc_geometry *__restrict__ g, uint32_t ts)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nne) return;
    const uint32_t s = order[N - nne + j];
    const SlotMeta m = meta[s];
    if (m.vol[0] < 0.f) return;   // never-filled slot
    const uint32_t T = g->nb_tables;
    uint32_t w = cells[top_index(*g, m.level, m.vol)];
    for (uint32_t i = 0; i + 1 < T; ++i) {
        if (word_state_of(w) != 1) return;
        const uint32_t b = w >> 2;
        tab.usage[i][b] = ts;
        w = tab.words[i][(uint64_t)b * table_b3(g->table[i]) + table_local(g->table[i], m.level, m.vol)];
    }
    if (word_state_of(w) != 1) return;
    const gc_table_geometry &t = g->table[T - 1];
    uint32_t *e = tab.words[T - 1] + (uint64_t)(w >> 2) * table_b3(t) + table_local(t, m.level, m.vol);
    if (*e == cell_word(s, 1)) *e = 0u;
}

// Serial state of one reference_new_bricks call.
struct RefCtx {
    uint4 *mrpd;
    uint32_t *cells;
    TableState tab;
    const gc_geometry *g;
    uint32_t ts;
    uint32_t used[GC_MAX_TABLES];   // blocks taken from each table's LRU tail
};

// the MRPD cell / table entry that maps the owner of block vb of table j
// (level and position of the brick it was taken for), if it still names vb;
// reads above table j-1 stamp their blocks (upward's get calls,
// CacheManagerKernels.hpp:246-263)
__device__ inline uint32_t *owner_word(RefCtx &c, uint32_t j, const SlotMeta &m, uint32_t vb, uint32_t *top)
{
    const gc_geometry &g = *c.g;
    *top = top_index(g, m.level, m.vol);
    uint32_t *p = c.cells + *top;
    for (uint32_t i = 0; i < j; ++i) {
        if (word_state_of(*p) != 1) return nullptr;
        const uint32_t b = *p >> 2;
        if (i + 1 < j) c.tab.usage[i][b] = c.ts;
        p = c.tab.words[i] + (uint64_t)b * table_b3(g.table[i]) + table_local(g.table[i], m.level, m.vol);
        *top = UINT32_MAX;
    }
    return *p == cell_word(vb, 1) ? p : nullptr;
}

// reference_new_bricks with downward / upward (CacheManagerKernels.hpp
// :148-357) for one brick at (level, vol): child is the brick's slot (w = 1)
// or 0 with w = 2 for an Empty brick.  downward walks the Mapped levels; the
// brick is referenced only when the walk ends at an Unmapped entry, and the
// tables below the last Mapped level each take their next LRU-tail block.
// upward, bottom table first: the block is cleared, its previous owner's
// entry cleared where it still names the block, the new entry written and
// the block's owner recorded; then the new chain is linked into the last
// Mapped level.  Departures from the reference's latent multi-table path,
// each one of its defects: its downward stack outlives the brick (a brick
// with an Unmapped MRPD entry is linked into an earlier brick's block,
// :322-350); fill_stack also takes a block from the table whose entry is
// Mapped and indexes lruSizes / indexLRU of the wrong table (:130-136,
// 176-183); an Empty brick is entered with w = 1 in its block and w = 2 in
// the MRPD, hiding the whole block (:210-214,226-232); the owner's entry is
// looked up one table too deep and cleared unchecked (:246-263).  Here a
// table with no block left in this load leaves the brick unreferenced.
__device__ void reference_one(RefCtx &c, uint32_t level, const float vol[3], uint32_t child, uint32_t cw,
                              uint32_t *unreferenced)
{
    const gc_geometry &g = *c.g;
    const uint32_t T = g.nb_tables;
    const uint32_t top = top_index(g, level, vol);
    uint32_t *p = c.cells + top;
    uint32_t d = 0;
    while (d < T && word_state_of(*p) == 1) {
        const uint32_t b = *p >> 2;
        c.tab.usage[d][b] = c.ts;
        p = c.tab.words[d] + (uint64_t)b * table_b3(g.table[d]) + table_local(g.table[d], level, vol);
        ++d;
    }
    if (word_state_of(*p) != 0) return;
    for (uint32_t j = d; j < T; ++j)
        if (c.used[j] >= g.table[j].nblocks) {
            ++*unreferenced;
            return;
        }
    uint32_t vic[GC_MAX_TABLES];
    for (uint32_t j = d; j < T; ++j) vic[j] = c.tab.order[j][g.table[j].nblocks - 1 - c.used[j]++];
    for (uint32_t j = T; j-- > d;) {
        const gc_table_geometry &t = g.table[j];
        const uint32_t b3 = table_b3(t);
        uint32_t *blk = c.tab.words[j] + (uint64_t)vic[j] * b3;
        for (uint32_t k = 0; k < b3; ++k) blk[k] = 0u;   // reset_entries (K_CacheDevice.hpp:279-311)
        SlotMeta &m = c.tab.meta[j][vic[j]];
        if (m.vol[0] > -1.f) {
            uint32_t otop;
            uint32_t *ow = owner_word(c, j, m, vic[j], &otop);
            if (ow) {
                *ow = 0u;
                if (otop != UINT32_MAX) c.mrpd[otop] = make_uint4(0, 0, 0, 0);
            }
        }
        blk[table_local(t, level, vol)] = j + 1 == T ? cell_word(child, cw) : cell_word(vic[j + 1], 1);
        m.vol[0] = vol[0];
        m.vol[1] = vol[1];
        m.vol[2] = vol[2];
        m.level = level;
    }
    if (d == T) {
        *p = cell_word(child, cw);
    } else {
        *p = cell_word(vic[d], 1);
        if (d == 0) {
            uint32_t o[3];
            block_origin(g.table[0], vic[0], o);
            c.mrpd[top] = make_uint4(o[0], o[1], o[2], 1);
        }
    }
}

__global__ void k_reference_tables(uint4 *mrpd, uint32_t *cells, TableState tab, const uint32_t *order, uint32_t N,
                                   uint32_t n, uint32_t nne, const uint32_t *asked4, const SlotMeta *meta,
                                   const gc_geometry *g, uint32_t ts, uint32_t *unreferenced, const uint32_t *gate)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (gate && *gate == 0u) return;   // the parallel referencing ran
    RefCtx c{mrpd, cells, tab, g, ts, {0u, 0u, 0u}};
    uint32_t unref = 0;
    // non-empty bricks (their slots already hold level and position), then
    // the Empty ones (CacheManagerKernels.hpp:336-353)
    for (uint32_t k = 0; k < nne; ++k) {
        const uint32_t s = order[N - nne + k];
        const SlotMeta m = meta[s];
        reference_one(c, m.level, m.vol, s, 1, &unref);
    }
    for (uint32_t k = nne; k < n; ++k) {
        const uint32_t *a = asked4 + 4 * k;
        float vol[3];
        volume_position(*g, a[3], a, vol);
        reference_one(c, a[3], vol, 0, 2, &unref);
    }
    *unreferenced = unref;
}

hipError_t launch_deref_tables(uint32_t *cells, const TableState &tab, const uint32_t *order, uint32_t nslots,
                               uint32_t nne, const SlotMeta *meta, const gc_geometry *d_geom, uint32_t ts,
                               hipStream_t stream)
{
    if (nne == 0) return hipSuccess;
    k_deref_tables<<<blocks_for(nne, BLK), BLK, 0, stream>>>(cells, tab, order, nslots, nne, meta, d_geom, ts);
    return hipGetLastError();
}

hipError_t launch_reference_tables(uint4 *mrpd, uint32_t *cells, const TableState &tab, const uint32_t *order,
                                   uint32_t nslots, uint32_t n, uint32_t nne, const uint32_t *asked4,
                                   const SlotMeta *meta, const gc_geometry *d_geom, uint32_t ts,
                                   uint32_t *unreferenced, hipStream_t stream)
{
    k_reference_tables<<<1, 1, 0, stream>>>(mrpd, cells, tab, order, nslots, n, nne, asked4, meta, d_geom, ts,
                                            unreferenced, nullptr);
    return hipGetLastError();
}

// ---------------------------------------------- parallel referencing, T = 1
// The serial referencing of one table cache, restated for the usual load:
// in brick order, a brick whose MRPD cell is Unmapped when it comes takes the
// next LRU-tail block (an event); the cell's later bricks find it Mapped.  A
// cell Unmapped at the start of the load stays so until its first brick,
// and a cell Mapped at the start stays Mapped unless its block is taken as a
// victim.  So if no cell the load touches is Mapped to one of the blocks the
// load takes, the events are exactly the first bricks of the cells Unmapped
// at the start, event i takes LRU position nblocks-1-i, the owner a victim
// loses is a cell the load does not touch, and every brick's entry is
// written independently.  k_ref_scan checks that condition; when it fails,
// these kernels change nothing and the serial kernel runs.

// LRU position of every table-0 block; clears the gate and the counter
__global__ void k_ref_init(const uint32_t *__restrict__ order, uint32_t nb, uint32_t *__restrict__ inv,
                           uint32_t *flag, uint32_t *unreferenced)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        flag[0] = 0u;
        *unreferenced = 0u;
    }
    if (i < nb) inv[order[i]] = i;
}

// the MRPD word of each brick's cell at the start of the load, and for a
// Mapped one its block's distance from the LRU tail
__global__ void k_ref_classify(const uint32_t *__restrict__ cells, RefScratch rs, uint32_t n, uint32_t nb)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t w = cells[rs.cell[k]];
    rs.w0[k] = w;
    rs.rank[k] = (w & 3u) == 1u ? nb - 1u - rs.inv[w >> 2] : UINT32_MAX;
}

// events in brick order (exclusive scan over first bricks of Unmapped
// cells) and the gate: 1 when a touched Mapped cell's block is among the
// min(events, nblocks) victims
__global__ __launch_bounds__(1024) void k_ref_scan(RefScratch rs, uint32_t n, uint32_t nb)
{
    __shared__ uint32_t part[1024];
    __shared__ uint32_t hit;
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n + 1023u) / 1024u, b0 = t * per, b1 = min(n, b0 + per);
    uint32_t c = 0;
    for (uint32_t k = b0; k < b1; ++k) c += (rs.rep[k] == k && (rs.w0[k] & 3u) == 0u) ? 1u : 0u;
    part[t] = c;
    if (t == 0) hit = 0u;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    const uint32_t E = part[1023];
    uint32_t e = part[t] - c;
    for (uint32_t k = b0; k < b1; ++k) {
        const bool ev = rs.rep[k] == k && (rs.w0[k] & 3u) == 0u;
        rs.ev[k] = ev ? e : UINT32_MAX;
        e += ev ? 1u : 0u;
    }
    const uint32_t taken = min(E, nb);
    bool h = false;
    for (uint32_t k = b0; k < b1; ++k) h = h || rs.rank[k] < taken;
    if (h) atomicOr(&hit, 1u);
    __syncthreads();
    if (t == 0) rs.flag[0] = hit;
}

// one workgroup per event brick: the victim block is cleared, its owner's
// MRPD cell cleared where it names the block, the brick's cell mapped to it
// and the block's owner recorded
__global__ __launch_bounds__(256) void k_ref_events(uint4 *__restrict__ mrpd, uint32_t *__restrict__ cells,
                                                    TableState tab, RefScratch rs, const uint32_t *__restrict__ order,
                                                    uint32_t N, uint32_t nne, const uint32_t *__restrict__ asked4,
                                                    const SlotMeta *__restrict__ meta, const gc_geometry *__restrict__ g)
{
    const uint32_t k = blockIdx.x;
    if (rs.flag[0]) return;
    const uint32_t i = rs.ev[k];
    const gc_table_geometry &t = g->table[0];
    if (i == UINT32_MAX || i >= t.nblocks) return;
    const uint32_t v = tab.order[0][t.nblocks - 1u - i];
    const uint32_t b3 = table_b3(t);
    uint32_t *blk = tab.words[0] + (uint64_t)v * b3;
    for (uint32_t e = threadIdx.x; e < b3; e += blockDim.x) blk[e] = 0u;   // reset_entries
    if (threadIdx.x != 0) return;
    SlotMeta &m = tab.meta[0][v];
    if (m.vol[0] > -1.f) {
        const uint32_t o = top_index(*g, m.level, m.vol);
        if (cells[o] == cell_word(v, 1)) {
            cells[o] = 0u;
            mrpd[o] = make_uint4(0, 0, 0, 0);
        }
    }
    SlotMeta nm;
    if (k < nne) {
        nm = meta[order[N - nne + k]];
    } else {
        const uint32_t *a = asked4 + 4 * k;
        nm.level = a[3];
        volume_position(*g, a[3], a, nm.vol);
    }
    m = nm;
    const uint32_t c = rs.cell[k];
    cells[c] = cell_word(v, 1);
    uint32_t o3[3];
    block_origin(t, v, o3);
    mrpd[c] = make_uint4(o3[0], o3[1], o3[2], 1);
}

// every brick's entry (downward's stamps of the blocks it walks through)
__global__ void k_ref_insert(TableState tab, RefScratch rs, const uint32_t *__restrict__ order, uint32_t N,
                             uint32_t n, uint32_t nne, const gc_geometry *__restrict__ g, uint32_t ts,
                             uint32_t *unreferenced)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || rs.flag[0]) return;
    const gc_table_geometry &t = g->table[0];
    const uint32_t r = rs.rep[k], w = rs.w0[r];
    uint32_t b;
    if ((w & 3u) == 1u) {
        b = w >> 2;
        tab.usage[0][b] = ts;
    } else if ((w & 3u) == 0u) {
        const uint32_t i = rs.ev[r];
        if (i >= t.nblocks) {
            atomicAdd(unreferenced, 1u);
            return;
        }
        b = tab.order[0][t.nblocks - 1u - i];
        if (k != r) tab.usage[0][b] = ts;
    } else {
        return;   // an Empty MRPD entry
    }
    uint32_t *e = tab.words[0] + (uint64_t)b * table_b3(t) + rs.loc[k];
    if ((*e & 3u) != 0u) return;
    *e = k < nne ? cell_word(order[N - nne + k], 1) : cell_word(0, 2);
}

hipError_t launch_reference_table1(uint4 *mrpd, uint32_t *cells, const TableState &tab, const uint32_t *order,
                                   uint32_t nslots, uint32_t n, uint32_t nne, const uint32_t *asked4,
                                   const SlotMeta *meta, const gc_geometry *d_geom, uint32_t ts,
                                   uint32_t *unreferenced, const RefScratch &rs, hipStream_t stream,
                                   uint32_t nblocks)
{
    if (n == 0) return hipSuccess;
    k_ref_init<<<blocks_for(nblocks, BLK), BLK, 0, stream>>>(tab.order[0], nblocks, rs.inv, rs.flag, unreferenced);
    k_ref_classify<<<blocks_for(n, BLK), BLK, 0, stream>>>(cells, rs, n, nblocks);
    k_ref_scan<<<1, 1024, 0, stream>>>(rs, n, nblocks);
    k_ref_events<<<n, 256, 0, stream>>>(mrpd, cells, tab, rs, order, nslots, nne, asked4, meta, d_geom);
    k_ref_insert<<<blocks_for(n, BLK), BLK, 0, stream>>>(tab, rs, order, nslots, n, nne, d_geom, ts, unreferenced);
    k_reference_tables<<<1, 1, 0, stream>>>(mrpd, cells, tab, order, nslots, n, nne, asked4, meta, d_geom, ts,
                                            unreferenced, rs.flag);
    return hipGetLastError();
}

hipError_t launch_occupancy(const SlotMeta *meta, uint32_t nslots, uint32_t *count, hipStream_t stream)
{
    hipError_t e = hipMemsetAsync(count, 0, sizeof(uint32_t), stream);
    if (e != hipSuccess) return e;
    if (nslots == 0) return hipSuccess;
    k_occupancy<<<blocks_for(nslots, BLK), BLK, 0, stream>>>(meta, nslots, count);
    return hipGetLastError();
}

hipError_t launch_mip(const void *in, const uint32_t dims[3], uint32_t bytes, void *out,
                      hipStream_t stream)
{
    const uint32_t OX = dims[0] / 2, OY = dims[1] / 2, OZ = dims[2] / 2;
    const uint64_t total = (uint64_t)OX * OY * OZ;
    if (total == 0) return hipSuccess;
    const uint32_t nb = (uint32_t)std::min<uint64_t>(blocks_for(total, BLK), 65536u * 4u);
    switch (bytes) {
    case 1: k_mip2<uint8_t><<<nb, BLK, 0, stream>>>((const uint8_t *)in, dims[0], dims[1], (uint8_t *)out, OX, OY, total); break;
    case 2: k_mip2<uint16_t><<<nb, BLK, 0, stream>>>((const uint16_t *)in, dims[0], dims[1], (uint16_t *)out, OX, OY, total); break;
    case 4: k_mip2<float><<<nb, BLK, 0, stream>>>((const float *)in, dims[0], dims[1], (float *)out, OX, OY, total); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_mip_brick(const MipBrickArgs &a, uint32_t bytes, const void *children, void *out,
                            hipStream_t stream)
{
    const uint32_t n = a.bs * a.bs * a.bs;
    const uint32_t nb = (n + BLK - 1) / BLK;
    switch (bytes) {
    case 1: k_mip_brick<uint8_t><<<nb, BLK, 0, stream>>>(a, (const uint8_t *)children, (uint8_t *)out); break;
    case 2: k_mip_brick<uint16_t><<<nb, BLK, 0, stream>>>(a, (const uint16_t *)children, (uint16_t *)out); break;
    case 4: k_mip_brick<float><<<nb, BLK, 0, stream>>>(a, (const float *)children, (float *)out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_cut_keys(const CutKeys &a, hipStream_t stream)
{
    if (a.n == 0) return hipSuccess;
    switch (a.bytes) {
    case 1: k_cut_keys<1><<<a.n, BLK, 0, stream>>>(a); break;
    case 2: k_cut_keys<2><<<a.n, BLK, 0, stream>>>(a); break;
    case 4: k_cut_keys<4><<<a.n, BLK, 0, stream>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_brick_cut(const void *level, const uint32_t dims[3], uint32_t bytes,
                            uint32_t bs, uint32_t cov, const uint32_t *bricks_xyz, uint32_t n,
                            void *out, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const uint8_t *l = static_cast<const uint8_t *>(level);
    uint8_t *o = static_cast<uint8_t *>(out);
    switch (bytes) {
    case 1: k_brick_cut<1><<<n, BLK, 0, stream>>>(l, dims[0], dims[1], dims[2], bs, cov, bricks_xyz, o); break;
    case 2: k_brick_cut<2><<<n, BLK, 0, stream>>>(l, dims[0], dims[1], dims[2], bs, cov, bricks_xyz, o); break;
    case 4: k_brick_cut<4><<<n, BLK, 0, stream>>>(l, dims[0], dims[1], dims[2], bs, cov, bricks_xyz, o); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} // namespace gc
