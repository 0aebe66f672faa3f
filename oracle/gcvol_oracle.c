/*
 * gcvol_oracle.c -- CPU restatement of GcCore's out-of-core volume path.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline).  Nothing in the
 * product library links, loads or calls this file.  See gcvol_oracle.h for
 * conventions; each function cites the reference rule it restates.
 *
 * Built with -O2 -ffp-contract=off (no FMA contraction, SSE single
 * precision), so every float expression below rounds exactly like the IEEE
 * single-precision expression written in the reference.
 */
#include "gcvol_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

/* ================================================================ geometry */

uint32_t or_level_count(const uint32_t dims[3], uint32_t brick_size)
{
    /* Mipmapper.cu:441-443 */
    const uint32_t max_dim = MAX(dims[0], MAX(dims[1], dims[2]));
    const uint32_t level_max =
        (uint32_t)ceil(log((double)(max_dim / brick_size)) / log(2.0));
    return level_max + 1;
}

void or_level_sizes(const uint32_t dims[3], const uint32_t ratio[3],
                    uint32_t nb_levels, uint32_t *out)
{
    /* Mipmapper.cu:583-597 */
    uint32_t s[3] = {dims[0], dims[1], dims[2]};
    for (uint32_t i = 0; i < nb_levels; ++i) {
        for (int a = 0; a < 3; ++a) {
            out[3 * i + a] = s[a];
            if (s[a] >= ratio[a]) s[a] /= ratio[a];
        }
    }
}

void or_init_meta(uint32_t nb_levels, const uint32_t *levels, uint32_t brick_size,
                  uint32_t covering, const uint32_t big_brick[3],
                  uint32_t *nb_bricks, uint32_t *nb_big_bricks, uint32_t *real_sizes)
{
    /* Bricker_v2.cpp:45-76: float ceil of the brick count, rounded up to a
     * multiple of the big-brick size. */
    const uint32_t inner = brick_size - 2 * covering;
    for (uint32_t i = 0; i < nb_levels; ++i) {
        for (int a = 0; a < 3; ++a) {
            uint32_t nb = (uint32_t)ceilf((float)levels[3 * i + a] / (float)inner);
            nb = (uint32_t)ceilf((float)nb / (float)big_brick[a]) * big_brick[a];
            if (nb_bricks) nb_bricks[3 * i + a] = nb;
            if (nb_big_bricks) nb_big_bricks[3 * i + a] = nb / big_brick[a];
            if (real_sizes) real_sizes[3 * i + a] = nb * inner;
        }
    }
}

void or_pool_shape(uint64_t budget_bytes, uint64_t brick_bytes, uint32_t out[3])
{
    /* Application.cu:85-93 */
    const uint64_t n = budget_bytes / brick_bytes;
    const uint64_t s = (uint64_t)cbrt((double)n);
    out[0] = out[1] = out[2] = (uint32_t)s;
    int d = 0;
    while ((uint64_t)out[0] * (uint64_t)out[1] * (uint64_t)out[2] < n) {
        out[d] += 1;
        d = (d + 1) % 3;
    }
}

void or_geometry(const or_cfg *cfg, or_geom *g)
{
    /* CacheManager.hpp:157-209 (the hierarchy: realNumberOfEntries bottom-up,
     * virtualizedNumberOfEntries top-down), MultiResolutionPageDirectory.hpp
     * :99-124, CacheDevice.hpp:123-180, CacheManager.hpp:223-285. */
    memset(g, 0, sizeof(*g));
    const uint32_t L = cfg->nb_levels, T = cfg->nb_tables;
    const uint32_t H = T + 2;   /* nbLevelPTHierarchy = CacheSize.size() + 1 */
    const double inner = (double)cfg->brick_size - 2.0 * (double)cfg->covering;
    /* real[h][l][a], virt[h][l][a] for hierarchy index h (0 = voxels) */
    uint32_t real[OR_MAX_TABLES + 2][OR_MAX_LEVELS][3], virt[OR_MAX_TABLES + 2][OR_MAX_LEVELS][3];
    for (uint32_t l = 0; l < L; ++l)
        for (int a = 0; a < 3; ++a) {
            real[0][l][a] = cfg->real[l][a];
            /* std::ceil(float(real0) / (bs - 2.0*c)) in double */
            real[1][l][a] = (uint32_t)ceil((double)(float)cfg->real[l][a] / inner);
            /* std::ceil(float(real[h-1]) / BlockSize[h-1]) in float */
            for (uint32_t h = 2; h < H; ++h)
                real[h][l][a] = (uint32_t)ceilf((float)real[h - 1][l][a] / (float)cfg->table_block[h - 2][a]);
            virt[H - 1][l][a] = real[H - 1][l][a];
            /* std::ceil(float(virt[h+1]) * BlockSize[h]) in float */
            for (uint32_t h = H - 2; h >= 1; --h)
                virt[h][l][a] = (uint32_t)ceilf((float)virt[h + 1][l][a] * (float)cfg->table_block[h - 1][a]);
            /* virtualized[0] = ceil(float(virtualized[1]) * (bs - 2.0*c)) */
            virt[0][l][a] = (uint32_t)ceil((double)(float)virt[1][l][a] * inner);
        }
    for (uint32_t l = 0; l < L; ++l) {
        for (int a = 0; a < 3; ++a) {
            g->level_dim[l][a] = real[1][l][a];
            g->virt_vox[l][a] = virt[0][l][a];
            g->rdv[l][a] = (float)cfg->real[l][a] / (float)g->virt_vox[l][a];
            g->ior[l][a] = (float)cfg->initial[l][a] / (float)cfg->real[l][a];
            g->top_dim[l][a] = real[H - 1][l][a];
        }
    }
    g->level_coord[0] = 0;
    g->top_coord[0] = 0;
    for (uint32_t l = 1; l < L; ++l) {
        g->level_coord[l] = g->level_coord[l - 1] + g->level_dim[l - 1][0];
        g->top_coord[l] = g->top_coord[l - 1] + g->top_dim[l - 1][0];
    }
    for (uint32_t l = 0; l < L; ++l) {
        g->mrpd_dim[0] += g->level_dim[l][0];
        g->mrpd_dim[1] = MAX(g->mrpd_dim[1], g->level_dim[l][1]);
        g->mrpd_dim[2] = MAX(g->mrpd_dim[2], g->level_dim[l][2]);
        g->top_cells_dim[0] += g->top_dim[l][0];
        g->top_cells_dim[1] = MAX(g->top_cells_dim[1], g->top_dim[l][1]);
        g->top_cells_dim[2] = MAX(g->top_cells_dim[2], g->top_dim[l][2]);
    }
    for (int a = 0; a < 3; ++a) {
        g->one_over[a] = 1.f / (float)(cfg->brick_size - 2 * cfg->covering);
        g->pool_dim[a] = cfg->cache_size[a] * cfg->brick_size;
    }
    g->nslots = cfg->cache_size[0] * cfg->cache_size[1] * cfg->cache_size[2];
    /* _tableCaches[i] = hierarchy index H - 2 - i (CacheManager.hpp:223-233) */
    g->nb_tables = T;
    for (uint32_t i = 0; i < T; ++i) {
        const uint32_t h = H - 2 - i;
        or_table_geom *t = &g->table[i];
        for (int a = 0; a < 3; ++a) {
            t->block[a] = cfg->table_block[h - 1][a];
            t->blocks[a] = cfg->table_cache[h - 1][a];
            t->surf[a] = t->block[a] * t->blocks[a];
            for (uint32_t l = 0; l < L; ++l) {
                t->virt[l][a] = virt[h][l][a];
                t->real[l][a] = real[h][l][a];
            }
        }
        t->nblocks = t->blocks[0] * t->blocks[1] * t->blocks[2];
    }
}

void or_volume_position(const or_geom *g, uint32_t level, const uint32_t brick[3], float out[3])
{
    /* CacheManagerKernels.cu:75-86 with epsilon = FLT_EPSILON
     * (CacheManager.hpp:554-561). */
    for (int a = 0; a < 3; ++a) {
        const float den = (float)g->virt_vox[level][a] * g->one_over[a];
        const float q = (float)brick[a] / den;
        out[a] = q + FLT_EPSILON;
    }
}

uint64_t or_mrpd_index(const or_geom *g, uint32_t level, const float pos[3])
{
    /* K_MultiResolutionPageDirectory.hpp:168-180 */
    const uint32_t x = g->level_coord[level] + (uint32_t)(pos[0] * (float)g->level_dim[level][0]);
    const uint32_t y = (uint32_t)(pos[1] * (float)g->level_dim[level][1]);
    const uint32_t z = (uint32_t)(pos[2] * (float)g->level_dim[level][2]);
    return (uint64_t)x + (uint64_t)g->mrpd_dim[0] * ((uint64_t)y + (uint64_t)g->mrpd_dim[1] * z);
}

uint64_t or_top_index(const or_geom *g, uint32_t level, const float pos[3])
{
    /* K_MultiResolutionPageDirectory.hpp:168-180 over the MRPD's levels */
    const uint32_t x = g->top_coord[level] + (uint32_t)(pos[0] * (float)g->top_dim[level][0]);
    const uint32_t y = (uint32_t)(pos[1] * (float)g->top_dim[level][1]);
    const uint32_t z = (uint32_t)(pos[2] * (float)g->top_dim[level][2]);
    return (uint64_t)x + (uint64_t)g->top_cells_dim[0] * ((uint64_t)y + (uint64_t)g->top_cells_dim[1] * z);
}

/* ============================================================ preprocessor */

void or_mip_reduce(const void *in, const uint32_t dims[3], uint32_t bytes,
                   const uint32_t ratio[3], void *out)
{
    /* Mipmapper.cu:30-105 over a whole level. */
    const uint32_t ox = dims[0] / ratio[0], oy = dims[1] / ratio[1], oz = dims[2] / ratio[2];
    const uint64_t X = dims[0], XY = (uint64_t)dims[0] * dims[1];
    const uint32_t area = ratio[0] * ratio[1] * ratio[2];
    for (uint32_t z = 0; z < oz; ++z)
        for (uint32_t y = 0; y < oy; ++y)
            for (uint32_t x = 0; x < ox; ++x) {
                const uint64_t o = x + (uint64_t)ox * (y + (uint64_t)oy * z);
                uint64_t isum = 0;
                double dsum = 0.0;
                for (uint32_t k = 0; k < ratio[2]; ++k)
                    for (uint32_t j = 0; j < ratio[1]; ++j)
                        for (uint32_t i = 0; i < ratio[0]; ++i) {
                            const uint64_t p = (x * ratio[0] + i) + X * (y * ratio[1] + j) +
                                               XY * (z * ratio[2] + k);
                            if (bytes == 1) isum += ((const uint8_t *)in)[p];
                            else if (bytes == 2) dsum += (double)((const uint16_t *)in)[p];
                            else dsum += (double)((const float *)in)[p];
                        }
                if (bytes == 1) ((uint8_t *)out)[o] = (uint8_t)(isum / area);
                else if (bytes == 2) ((uint16_t *)out)[o] = (uint16_t)(dsum / (double)area);
                else ((float *)out)[o] = (float)(dsum / (double)area);
            }
}

static inline void put_voxel(void *dst, uint64_t i, uint32_t bytes, const void *src, uint64_t j)
{
    memcpy((uint8_t *)dst + i * bytes, (const uint8_t *)src + j * bytes, bytes);
}

void or_brick_extract(const void *level, const uint32_t dims[3], uint32_t bytes,
                      uint32_t bs, uint32_t cov, const uint32_t bb[3],
                      const uint32_t bbc[3], void *out)
{
    /* Bricker_v2.cpp:138-222 and get_iterator_from_indexes (:295-299). */
    const uint32_t inner = bs - 2 * cov;
    const uint64_t BX = (uint64_t)bs * bb[0], BY = (uint64_t)bs * bb[1], BZ = (uint64_t)bs * bb[2];
    memset(out, 0, BX * BY * BZ * bytes);
    for (uint32_t sz = 0; sz < bb[2]; ++sz)
        for (uint32_t sy = 0; sy < bb[1]; ++sy)
            for (uint32_t sx = 0; sx < bb[0]; ++sx) {
                const int64_t ox = (int64_t)(bbc[0] * bb[0] + sx) * inner - cov;
                const int64_t oy = (int64_t)(bbc[1] * bb[1] + sy) * inner - cov;
                const int64_t oz = (int64_t)(bbc[2] * bb[2] + sz) * inner - cov;
                for (uint32_t k = 0; k < bs; ++k) {
                    const int64_t gz = oz + k;
                    if (gz < 0 || gz >= dims[2]) continue;
                    for (uint32_t j = 0; j < bs; ++j) {
                        const int64_t gy = oy + j;
                        if (gy < 0 || gy >= dims[1]) continue;
                        for (uint32_t i = 0; i < bs; ++i) {
                            const int64_t gx = ox + i;
                            if (gx < 0 || gx >= dims[0]) continue;
                            const uint64_t d = (sx * (uint64_t)bs + i) +
                                               BX * ((sy * (uint64_t)bs + j) + BY * (sz * (uint64_t)bs + k));
                            const uint64_t s = (uint64_t)gx +
                                               (uint64_t)dims[0] * ((uint64_t)gy + (uint64_t)dims[1] * gz);
                            put_voxel(out, d, bytes, level, s);
                        }
                    }
                }
            }
}

void or_brick_level(const void *level, const uint32_t dims[3], uint32_t bytes, uint32_t bs, uint32_t cov,
                    uint32_t b0, uint32_t b1, void *out)
{
    /* process_bricking's per-level loop (Bricker_v2.cpp:79-252) over bricks
     * [b0, b1) of the level in x-fastest brick order, one brick of bs^3
     * voxels after the other in `out`. */
    const uint32_t inner = bs - 2 * cov;
    const uint32_t nb[3] = {(dims[0] + inner - 1) / inner, (dims[1] + inner - 1) / inner, (dims[2] + inner - 1) / inner};
    const uint32_t one[3] = {1, 1, 1};
    const uint64_t bvox = (uint64_t)bs * bs * bs;
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t c[3] = {b % nb[0], (b / nb[0]) % nb[1], b / (nb[0] * nb[1])};
        or_brick_extract(level, dims, bytes, bs, cov, one, c, (uint8_t *)out + (uint64_t)(b - b0) * bvox * bytes);
    }
}

/* =================================================================== cache */

struct or_cache {
    or_cfg cfg;
    or_geom g;
    const void *levels[OR_MAX_LEVELS];
    uint32_t *mrpd, *req, *usage;
    or_lru *lru, *lru_tmp;
    uint8_t *pool;
    uint64_t mrpd_cells;      /* brick grid (request buffer) */
    uint64_t top_cells;       /* the MRPD                   */
    /* table caches (CacheDevice<uint4>): entries, block usage, LRU */
    uint32_t *tab[OR_MAX_TABLES], *tab_usage[OR_MAX_TABLES];
    or_lru *tab_lru[OR_MAX_TABLES], *tab_lru_tmp[OR_MAX_TABLES];
    uint32_t unreferenced;
    uint32_t ts;
    /* request handler state (RequestHandler.hpp:89-105) */
    int request_done;
    uint32_t *asked;          /* n x uint4 */
    uint32_t nb_asked, nb_non_empty;
    uint32_t *bb_indexes;     /* per non-empty brick: index of its big brick */
    uint32_t *bb_coords;      /* per big brick: uint4 */
    uint8_t *staging;
};

const or_geom *or_cache_geom(const or_cache *c) { return &c->g; }
uint32_t *or_cache_mrpd(or_cache *c) { return c->mrpd; }
uint32_t *or_cache_request(or_cache *c) { return c->req; }
uint32_t *or_cache_usage(or_cache *c) { return c->usage; }
or_lru *or_cache_lru(or_cache *c) { return c->lru; }
void *or_cache_pool(or_cache *c) { return c->pool; }
uint32_t or_cache_timestamp(const or_cache *c) { return c->ts; }
uint32_t *or_cache_table(or_cache *c, uint32_t i) { return i < c->g.nb_tables ? c->tab[i] : NULL; }
uint32_t *or_cache_table_usage(or_cache *c, uint32_t i) { return i < c->g.nb_tables ? c->tab_usage[i] : NULL; }
or_lru *or_cache_table_lru(or_cache *c, uint32_t i) { return i < c->g.nb_tables ? c->tab_lru[i] : NULL; }
uint32_t or_cache_unreferenced(const or_cache *c) { return c->unreferenced; }

static void lru_init(or_lru *lru, const uint32_t n[3], const uint32_t block[3])
{
    /* LRU initialisation, x fastest (CacheDevice.hpp:151-165) */
    for (uint32_t z = 0; z < n[2]; ++z)
        for (uint32_t y = 0; y < n[1]; ++y)
            for (uint32_t x = 0; x < n[0]; ++x) {
                or_lru *e = &lru[x + y * n[0] + z * n[0] * n[1]];
                e->cache_pos[0] = x * block[0];
                e->cache_pos[1] = y * block[1];
                e->cache_pos[2] = z * block[2];
                e->level = 0;
                e->vol_pos[0] = e->vol_pos[1] = e->vol_pos[2] = -1.f;
            }
}

or_cache *or_cache_create(const or_cfg *cfg, const void *const *levels)
{
    or_cache *c = (or_cache *)calloc(1, sizeof(or_cache));
    c->cfg = *cfg;
    or_geometry(cfg, &c->g);
    for (uint32_t l = 0; l < cfg->nb_levels; ++l) c->levels[l] = levels ? levels[l] : NULL;
    c->mrpd_cells = (uint64_t)c->g.mrpd_dim[0] * c->g.mrpd_dim[1] * c->g.mrpd_dim[2];
    c->top_cells = (uint64_t)c->g.top_cells_dim[0] * c->g.top_cells_dim[1] * c->g.top_cells_dim[2];
    c->mrpd = (uint32_t *)calloc(c->top_cells * 4, sizeof(uint32_t));
    c->req = (uint32_t *)calloc(c->mrpd_cells, sizeof(uint32_t));
    c->usage = (uint32_t *)calloc(c->g.nslots, sizeof(uint32_t));
    c->lru = (or_lru *)calloc(c->g.nslots, sizeof(or_lru));
    c->lru_tmp = (or_lru *)calloc(c->g.nslots, sizeof(or_lru));
    const uint64_t pool_vox = (uint64_t)c->g.pool_dim[0] * c->g.pool_dim[1] * c->g.pool_dim[2];
    c->pool = (uint8_t *)calloc(pool_vox, cfg->bytes);
    const uint32_t bs3[3] = {cfg->brick_size, cfg->brick_size, cfg->brick_size};
    lru_init(c->lru, cfg->cache_size, bs3);
    for (uint32_t i = 0; i < c->g.nb_tables; ++i) {
        const or_table_geom *t = &c->g.table[i];
        const uint64_t ne = (uint64_t)t->surf[0] * t->surf[1] * t->surf[2];
        c->tab[i] = (uint32_t *)calloc(ne * 4, sizeof(uint32_t));
        c->tab_usage[i] = (uint32_t *)calloc(t->nblocks, sizeof(uint32_t));
        c->tab_lru[i] = (or_lru *)calloc(t->nblocks, sizeof(or_lru));
        c->tab_lru_tmp[i] = (or_lru *)calloc(t->nblocks, sizeof(or_lru));
        lru_init(c->tab_lru[i], t->blocks, t->block);
    }
    c->ts = 1; /* CacheManager.hpp:152 */
    return c;
}

void or_cache_destroy(or_cache *c)
{
    if (!c) return;
    free(c->mrpd); free(c->req); free(c->usage); free(c->lru); free(c->lru_tmp);
    free(c->pool); free(c->asked); free(c->bb_indexes); free(c->bb_coords); free(c->staging);
    for (uint32_t i = 0; i < c->g.nb_tables; ++i) {
        free(c->tab[i]); free(c->tab_usage[i]); free(c->tab_lru[i]); free(c->tab_lru_tmp[i]);
    }
    free(c);
}

static void lru_partition(or_lru **lru, or_lru **tmp, uint32_t n, const uint32_t *usage,
                          const uint32_t blocks[3], const uint32_t block[3], uint32_t ts)
{
    /* create_usage_mask (CacheManagerKernels.cu:8-29) then the two copy_if of
     * CacheDevice::update (CacheDevice.hpp:202-219): used entries first, both
     * halves in their previous order. */
    or_lru *a = *lru, *b = *tmp;
    uint32_t k = 0;
    for (int pass = 0; pass < 2; ++pass)
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t *cp = a[i].cache_pos;
            const uint64_t u = (uint64_t)(cp[0] / block[0]) +
                               (uint64_t)blocks[0] * ((cp[1] / block[1]) + (uint64_t)blocks[1] * (cp[2] / block[2]));
            if ((usage[u] == ts) == (pass == 0)) b[k++] = a[i];
        }
    *lru = b;
    *tmp = a;
}

void or_lru_update(or_cache *c)
{
    /* CacheManager.hpp:387-406: the data cache, then each table cache */
    const uint32_t bs3[3] = {c->cfg.brick_size, c->cfg.brick_size, c->cfg.brick_size};
    lru_partition(&c->lru, &c->lru_tmp, c->g.nslots, c->usage, c->cfg.cache_size, bs3, c->ts);
    for (uint32_t i = 0; i < c->g.nb_tables; ++i) {
        const or_table_geom *t = &c->g.table[i];
        lru_partition(&c->tab_lru[i], &c->tab_lru_tmp[i], t->nblocks, c->tab_usage[i], t->blocks, t->block, c->ts);
    }
}

uint32_t or_request_list(or_cache *c, uint32_t *out4, uint32_t cap)
{
    /* thrust::copy_if over the counting range [0, size-1) -- the last cell is
     * never collected -- then min(count, cap) entries through coord_1D_to_3D
     * (CacheManager.hpp:410-474, CacheManagerKernels.cu:32-72). */
    const uint64_t N = c->mrpd_cells;
    uint32_t n = 0;
    uint32_t sx = 0;
    for (uint32_t l = 0; l < c->cfg.nb_levels; ++l) sx += c->g.level_dim[l][0];
    const uint64_t SX = sx, SY = c->g.level_dim[0][1];
    for (uint64_t i = 0; i + 1 < N && n < cap; ++i) {
        if (c->req[i] != c->ts) continue;
        uint64_t idx = i;
        const uint32_t z = (uint32_t)(idx / (SX * SY));
        idx -= (uint64_t)z * (SX * SY);
        const uint32_t y = (uint32_t)(idx / SX);
        idx -= (uint64_t)y * SX;
        const uint32_t x = (uint32_t)idx;
        uint32_t sum = 0, level;
        for (level = 0; level < c->cfg.nb_levels; ++level) {
            sum += c->g.level_dim[level][0];
            if (x < sum) break;
        }
        out4[4 * n + 0] = x - (sum - c->g.level_dim[level][0]);
        out4[4 * n + 1] = y;
        out4[4 * n + 2] = z;
        out4[4 * n + 3] = level;
        ++n;
    }
    return n;
}

static int big_brick_empty(const or_cache *c, const uint8_t *bb, uint64_t nvox)
{
    const float thr = c->cfg.empty_threshold;
    for (uint64_t i = 0; i < nvox; ++i) {
        float v;
        if (c->cfg.bytes == 1) v = (float)bb[i];
        else if (c->cfg.bytes == 2) v = (float)((const uint16_t *)bb)[i];
        else v = ((const float *)bb)[i];
        if (!(v < thr)) return 0;
    }
    return 1;
}

/* uint4 lexicographic order (Operator.hpp:3-22) for the std::map of
 * RequestHandler.hpp:289; a sorted array with binary search is equivalent. */
typedef struct { uint32_t k[4]; int32_t v; } bb_entry;

static int bb_cmp(const uint32_t *a, const uint32_t *b)
{
    for (int i = 0; i < 4; ++i) {
        if (a[i] < b[i]) return -1;
        if (a[i] > b[i]) return 1;
    }
    return 0;
}

static void handle_request(or_cache *c, const uint32_t *asked, uint32_t n)
{
    /* RequestHandler.hpp:279-372 */
    const uint32_t *BB = c->cfg.big_brick;
    const uint32_t bs = c->cfg.brick_size, bytes = c->cfg.bytes;
    const uint64_t bb_vox = (uint64_t)bs * bs * bs * BB[0] * BB[1] * BB[2];
    free(c->asked); free(c->bb_indexes); free(c->bb_coords); free(c->staging);
    c->asked = (uint32_t *)malloc((size_t)n * 4 * sizeof(uint32_t));
    c->bb_indexes = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    c->bb_coords = (uint32_t *)malloc((size_t)n * 4 * sizeof(uint32_t));
    c->staging = (uint8_t *)malloc((size_t)(n ? n : 1) * bb_vox * bytes);
    uint8_t *empty = (uint8_t *)calloc(n ? n : 1, 1);
    bb_entry *map = (bb_entry *)malloc((size_t)(n ? n : 1) * sizeof(bb_entry));
    uint32_t map_n = 0, cpt = 0, nne = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t *a = asked + 4 * i;
        const uint32_t key[4] = {a[0] / BB[0], a[1] / BB[1], a[2] / BB[2], a[3]};
        uint32_t lo = 0, hi = map_n;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (bb_cmp(map[mid].k, key) < 0) lo = mid + 1; else hi = mid;
        }
        if (lo < map_n && bb_cmp(map[lo].k, key) == 0) {
            if (map[lo].v == -1) empty[i] = 1;
            else { empty[i] = 0; c->bb_indexes[nne++] = (uint32_t)map[lo].v; }
            continue;
        }
        const uint32_t level = key[3];
        uint8_t *dst = c->staging + (uint64_t)cpt * bb_vox * bytes;
        const uint32_t dims[3] = {c->cfg.initial[level][0], c->cfg.initial[level][1], c->cfg.initial[level][2]};
        or_brick_extract(c->levels[level], dims, bytes, bs, c->cfg.covering, BB, key, dst);
        const int is_empty = big_brick_empty(c, dst, bb_vox);
        memmove(map + lo + 1, map + lo, (size_t)(map_n - lo) * sizeof(bb_entry));
        memcpy(map[lo].k, key, sizeof(key));
        ++map_n;
        if (!is_empty) {
            map[lo].v = (int32_t)cpt;
            c->bb_indexes[nne] = cpt;
            memcpy(c->bb_coords + 4 * cpt, key, sizeof(key));
            ++cpt;
            empty[i] = 0;
            ++nne;
        } else {
            map[lo].v = -1;
            empty[i] = 1;
        }
    }
    /* stable partition: non-empty first (RequestHandler.hpp:354-369) */
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (!empty[i]) memcpy(c->asked + 4 * k++, asked + 4 * i, 4 * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i)
        if (empty[i]) memcpy(c->asked + 4 * k++, asked + 4 * i, 4 * sizeof(uint32_t));
    c->nb_asked = n;
    c->nb_non_empty = nne;
    free(empty);
    free(map);
}

uint32_t or_asked(const or_cache *c, uint32_t *out4, uint32_t *nb_non_empty)
{
    if (out4 && c->nb_asked) memcpy(out4, c->asked, (size_t)c->nb_asked * 4 * sizeof(uint32_t));
    if (nb_non_empty) *nb_non_empty = c->nb_non_empty;
    return c->nb_asked;
}

static void mrpd_set(or_cache *c, uint32_t level, const float pos[3], const uint32_t v[4])
{
    memcpy(c->mrpd + 4 * or_mrpd_index(&c->g, level, pos), v, 4 * sizeof(uint32_t));
}

/* ------------------------------------------------ page-table caches (tables)
 *
 * The hierarchy MRPD -> table 0 -> ... -> table T-1 -> data cache of
 * CacheManager.hpp:151-236.  An entry is a uint4 (x, y, z, w): w = 0
 * Unmapped, 1 Mapped, 2 Empty; x, y, z the origin of the block (or brick
 * slot) it maps in the next cache's surface.  The entry of a position inside
 * a block at origin o is o + (uint)(pos * virtualized) % BlockSize
 * (K_CacheDevice.hpp:221-233,316-334, covering 0). */

static inline uint64_t tab_entry(const or_table_geom *t, const uint32_t o[3], uint32_t level, const float pos[3])
{
    uint32_t e[3];
    for (int a = 0; a < 3; ++a) e[a] = o[a] + ((uint32_t)(pos[a] * (float)t->virt[level][a])) % t->block[a];
    return (uint64_t)e[0] + (uint64_t)t->surf[0] * ((uint64_t)e[1] + (uint64_t)t->surf[1] * e[2]);
}

/* update_usage_buffer of a table block (K_CacheDevice.hpp:245-250,323-327) */
static inline void tab_stamp(or_cache *c, uint32_t i, const uint32_t o[3], uint32_t ts)
{
    const or_table_geom *t = &c->g.table[i];
    const uint64_t u = (uint64_t)(o[0] / t->block[0]) +
                       (uint64_t)t->blocks[0] * ((o[1] / t->block[1]) + (uint64_t)t->blocks[1] * (o[2] / t->block[2]));
    __atomic_store_n(&c->tab_usage[i][u], ts, __ATOMIC_RELAXED);
}

/* The page-table walk of get / get_normalized (K_CacheManager.hpp:206-231,
 * 310-339): the MRPD entry, then one entry per table cache while the entries
 * are Mapped; each table read stamps its block's usage when stamp != 0.  e
 * receives the last entry read: w = 1 maps the data brick, 0 Unmapped and 2
 * Empty end the walk at whichever level they occur.  The reference keeps
 * walking after an Unmapped table entry with a coarser level pending and then
 * repeats the same lookup for ever (checkOtherLevel without ++level,
 * K_CacheManager.hpp:212-222,318-330); here such an entry is Unmapped, as an
 * Unmapped MRPD entry is. */
static void walk(or_cache *c, uint32_t level, const float pv[3], int stamp, uint32_t ts, uint32_t e[4])
{
    memcpy(e, c->mrpd + 4 * or_top_index(&c->g, level, pv), 4 * sizeof(uint32_t));
    for (uint32_t i = 0; i < c->g.nb_tables && e[3] == 1; ++i) {
        if (stamp) tab_stamp(c, i, e, ts);
        memcpy(e, c->tab[i] + 4 * tab_entry(&c->g.table[i], e, level, pv), 4 * sizeof(uint32_t));
    }
}

/* The key of the entries a walk reads: the MRPD cell, then per table cache
 * the entry's index inside its block.  Two positions with the same key walk
 * to the same entry. */
static inline void walk_key(const or_geom *g, uint32_t level, const float pv[3], uint64_t key[1 + OR_MAX_TABLES])
{
    key[0] = or_top_index(g, level, pv);
    for (uint32_t i = 0; i < OR_MAX_TABLES; ++i) {
        key[1 + i] = 0;
        if (i >= g->nb_tables) continue;
        const or_table_geom *t = &g->table[i];
        uint32_t l[3];
        for (int a = 0; a < 3; ++a) l[a] = ((uint32_t)(pv[a] * (float)t->virt[level][a])) % t->block[a];
        key[1 + i] = l[0] + (uint64_t)t->block[0] * (l[1] + (uint64_t)t->block[1] * l[2]);
    }
}

static const uint32_t zero4[4] = {0, 0, 0, 0};

/* dereference_old_bricks (CacheManagerKernels.hpp:111-145) for a victim slot
 * of the data cache: the entry that maps it is cleared.  The reference
 * clears whatever entry the walk reaches (for a table walk even from an
 * Unmapped entry, at the origin (0,0,0)); once the table block that held the
 * slot's entry was evicted, that entry belongs to another brick, whose
 * mapping it would drop.  Here the walk stops at an entry that is not
 * Mapped and clears the last entry only if it still names this slot. */
static void deref_tables(or_cache *c, const or_lru *v)
{
    const uint32_t T = c->g.nb_tables;
    uint32_t e[4];
    memcpy(e, c->mrpd + 4 * or_top_index(&c->g, v->level, v->vol_pos), sizeof(e));
    for (uint32_t i = 0; i + 1 < T; ++i) {
        if (e[3] != 1) return;
        tab_stamp(c, i, e, c->ts);
        memcpy(e, c->tab[i] + 4 * tab_entry(&c->g.table[i], e, v->level, v->vol_pos), sizeof(e));
    }
    if (e[3] != 1) return;
    uint32_t *last = c->tab[T - 1] + 4 * tab_entry(&c->g.table[T - 1], e, v->level, v->vol_pos);
    if (last[3] == 1 && last[0] == v->cache_pos[0] && last[1] == v->cache_pos[1] && last[2] == v->cache_pos[2])
        memcpy(last, zero4, sizeof(zero4));
}

/* The entry that maps block `o` of table j for the position (level, pos) it
 * was assigned to: the MRPD cell for j = 0, else the entry in the table j-1
 * block the walk reaches; NULL if the walk meets an entry that is not Mapped
 * or the entry no longer names o.  Table reads above table j-1 stamp usage
 * (the get calls of upward, CacheManagerKernels.hpp:246-263). */
static uint32_t *owner_entry(or_cache *c, uint32_t j, uint32_t level, const float pos[3], const uint32_t o[3])
{
    uint32_t *p = c->mrpd + 4 * or_top_index(&c->g, level, pos);
    for (uint32_t i = 0; i < j; ++i) {
        if (p[3] != 1) return NULL;
        uint32_t e[3] = {p[0], p[1], p[2]};
        if (i + 1 < j) tab_stamp(c, i, e, c->ts);
        p = c->tab[i] + 4 * tab_entry(&c->g.table[i], e, level, pos);
    }
    if (p[3] != 1 || p[0] != o[0] || p[1] != o[1] || p[2] != o[2]) return NULL;
    return p;
}

/* reference_new_bricks with downward / upward (CacheManagerKernels.hpp
 * :148-357) for one brick: `child` is the brick's slot origin (w = 1) or, for
 * an Empty brick, (0,0,0) with w = 2; used[j] counts the blocks of table j
 * taken from its LRU tail in this load.
 *
 *   downward: walk the Mapped levels; the brick is referenced only if the
 *     walk ends at an Unmapped entry (the MRPD-only rule "insert if
 *     Unmapped", :205-208); the tables below the last Mapped level each take
 *     their next LRU-tail block (fill_stack, :128-136).
 *   upward, bottom table first: the block is cleared (reset_entries), the
 *     entry of its previous owner is cleared, the new entry written and the
 *     block's LRU record takes the brick's level and position (:236-301);
 *     then the new chain is linked into the last Mapped level.
 *
 * Departures from the reference, each a defect of its latent multi-table
 * path: the stack of downward is kept across bricks, so a brick whose MRPD
 * entry is Unmapped is linked into the block of an earlier brick (:322-350);
 * fill_stack takes a block also from the table whose entry is already
 * Mapped (:176-183) and indexes lruSizes / indexLRU of the wrong table
 * (:130-136); an Empty brick is entered with w = 1 in its table block and
 * w = 2 in the MRPD, which hides the whole block (:210-214,226-232); the
 * previous owner's entry is looked up one table too deep (:246-263) and
 * cleared without checking that it still names the block.  Here only
 * missing blocks are taken, Empty is entered as w = 2 at the brick's own
 * entry, owners are cleared only where they still name the block, and a
 * brick for whose tables this load has no block left stays unreferenced. */
static void reference_one(or_cache *c, uint32_t level, const float pos[3], const uint32_t child[3], uint32_t cw,
                          uint32_t used[OR_MAX_TABLES])
{
    const uint32_t T = c->g.nb_tables;
    uint32_t *p = c->mrpd + 4 * or_top_index(&c->g, level, pos);
    uint32_t d = 0;
    while (d < T && p[3] == 1) {
        const uint32_t o[3] = {p[0], p[1], p[2]};
        tab_stamp(c, d, o, c->ts);
        p = c->tab[d] + 4 * tab_entry(&c->g.table[d], o, level, pos);
        ++d;
    }
    if (p[3] != 0) return;
    for (uint32_t j = d; j < T; ++j)
        if (used[j] >= c->g.table[j].nblocks) {
            ++c->unreferenced;
            return;
        }
    or_lru *vic[OR_MAX_TABLES];
    for (uint32_t j = d; j < T; ++j) vic[j] = &c->tab_lru[j][c->g.table[j].nblocks - 1 - used[j]++];
    for (uint32_t j = T; j-- > d;) {
        const or_table_geom *t = &c->g.table[j];
        or_lru *v = vic[j];
        for (uint32_t z = 0; z < t->block[2]; ++z)
            for (uint32_t y = 0; y < t->block[1]; ++y)
                for (uint32_t x = 0; x < t->block[0]; ++x) {
                    const uint64_t k = (uint64_t)(v->cache_pos[0] + x) +
                                       (uint64_t)t->surf[0] * ((v->cache_pos[1] + y) + (uint64_t)t->surf[1] * (v->cache_pos[2] + z));
                    memcpy(c->tab[j] + 4 * k, zero4, sizeof(zero4));
                }
        if (v->vol_pos[0] > -1.f) {
            uint32_t *owner = owner_entry(c, j, v->level, v->vol_pos, v->cache_pos);
            if (owner) memcpy(owner, zero4, sizeof(zero4));
        }
        uint32_t *dst = c->tab[j] + 4 * tab_entry(t, v->cache_pos, level, pos);
        if (j + 1 == T) {
            dst[0] = child[0]; dst[1] = child[1]; dst[2] = child[2]; dst[3] = cw;
        } else {
            dst[0] = vic[j + 1]->cache_pos[0]; dst[1] = vic[j + 1]->cache_pos[1]; dst[2] = vic[j + 1]->cache_pos[2];
            dst[3] = 1;
        }
        memcpy(v->vol_pos, pos, 3 * sizeof(float));
        v->level = level;
    }
    if (d == T) {
        p[0] = child[0]; p[1] = child[1]; p[2] = child[2]; p[3] = cw;
    } else {
        p[0] = vic[d]->cache_pos[0]; p[1] = vic[d]->cache_pos[1]; p[2] = vic[d]->cache_pos[2]; p[3] = 1;
    }
}

static void load_new_bricks(or_cache *c)
{
    /* CacheManager.hpp:503-632 with CacheManagerKernels.hpp:59-145,308-357
     * and CacheManagerKernels.cu:75-114 (MRPD-only hierarchy). */
    const uint32_t n = c->nb_asked;
    if (n == 0) return;
    const uint32_t nne = c->nb_non_empty;
    const uint32_t N = c->g.nslots;
    or_lru *victims = c->lru + (N - nne);
    /* dereference_old_bricks: victims that were mapped lose their entry */
    for (uint32_t j = 0; j < nne; ++j) {
        if (victims[j].vol_pos[0] < 0.f) continue;
        if (c->g.nb_tables) deref_tables(c, &victims[j]);
        else mrpd_set(c, victims[j].level, victims[j].vol_pos, zero4);
    }
    /* compute_lru_volume_position into a zero-initialised draft */
    or_lru *draft = (or_lru *)calloc(n, sizeof(or_lru));
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t level = c->asked[4 * k + 3];
        or_volume_position(&c->g, level, c->asked + 4 * k, draft[k].vol_pos);
        draft[k].level = level;
    }
    /* copy_lru_position + write_brick */
    const uint32_t bs = c->cfg.brick_size, bytes = c->cfg.bytes;
    const uint32_t *BB = c->cfg.big_brick;
    const uint64_t BVX = (uint64_t)BB[0] * bs, BVY = (uint64_t)BB[1] * bs, BVZ = (uint64_t)BB[2] * bs;
    const uint64_t bb_vox = BVX * BVY * BVZ;
    const uint64_t PX = c->g.pool_dim[0], PXY = (uint64_t)c->g.pool_dim[0] * c->g.pool_dim[1];
    for (uint32_t j = 0; j < nne; ++j) {
        victims[j].level = draft[j].level;
        memcpy(victims[j].vol_pos, draft[j].vol_pos, sizeof(draft[j].vol_pos));
        const uint32_t bbi = c->bb_indexes[j];
        const uint32_t *sb = c->asked + 4 * j;
        const uint32_t *bbc = c->bb_coords + 4 * bbi;
        const uint64_t ox = sb[0] - bbc[0] * BB[0], oy = sb[1] - bbc[1] * BB[1], oz = sb[2] - bbc[2] * BB[2];
        const uint64_t base = bbi * bb_vox + ((oz * BVY + oy) * BVX + ox) * bs;
        for (uint32_t z = 0; z < bs; ++z)
            for (uint32_t y = 0; y < bs; ++y)
                for (uint32_t x = 0; x < bs; ++x) {
                    const uint64_t src = base + (z * BVY + y) * BVX + x;
                    const uint64_t dst = (victims[j].cache_pos[0] + x) + PX * (victims[j].cache_pos[1] + y) +
                                         PXY * (victims[j].cache_pos[2] + z);
                    put_voxel(c->pool, dst, bytes, c->staging, src);
                }
    }
    c->unreferenced = 0;
    if (c->g.nb_tables) {
        /* reference_new_bricks through the table caches: non-empty bricks,
         * then the empty ones, one at a time (CacheManagerKernels.hpp:336-353) */
        uint32_t used[OR_MAX_TABLES] = {0, 0, 0};
        static const uint32_t origin0[3] = {0, 0, 0};
        for (uint32_t j = 0; j < nne; ++j)
            reference_one(c, victims[j].level, victims[j].vol_pos, victims[j].cache_pos, 1, used);
        for (uint32_t k = nne; k < n; ++k) reference_one(c, draft[k].level, draft[k].vol_pos, origin0, 2, used);
        free(draft);
        return;
    }
    /* reference_new_bricks (MRPD only): insert if the entry is Unmapped */
    for (uint32_t j = 0; j < nne; ++j) {
        const uint64_t idx = or_mrpd_index(&c->g, victims[j].level, victims[j].vol_pos);
        if (c->mrpd[4 * idx + 3] >= 1) continue;
        const uint32_t v[4] = {victims[j].cache_pos[0], victims[j].cache_pos[1], victims[j].cache_pos[2], 1};
        memcpy(c->mrpd + 4 * idx, v, sizeof(v));
    }
    for (uint32_t k = nne; k < n; ++k) {
        const uint64_t idx = or_mrpd_index(&c->g, draft[k].level, draft[k].vol_pos);
        if (c->mrpd[4 * idx + 3] >= 1) continue;
        const uint32_t v[4] = {draft[k].cache_pos[0], draft[k].cache_pos[1], draft[k].cache_pos[2], 2};
        memcpy(c->mrpd + 4 * idx, v, sizeof(v));
    }
    free(draft);
}

int or_cache_update(or_cache *c)
{
    /* CacheManager.hpp:319-354 with a handler that completes inside
     * notify_request, so is_working() is false on every call. */
    or_lru_update(c);
    int check;
    if (c->request_done) {
        load_new_bricks(c);
        c->request_done = 0;
        check = 0;
    } else {
        const uint32_t cap = c->cfg.max_requests;
        uint32_t *buf = (uint32_t *)malloc((size_t)(cap ? cap : 1) * 4 * sizeof(uint32_t));
        const uint32_t n = or_request_list(c, buf, cap);
        if (n == 0) {
            check = 1; /* prepare_bricks: nothing requested (:486) */
        } else {
            handle_request(c, buf, n);
            c->request_done = 1;
            check = 0;
        }
        free(buf);
    }
    ++c->ts;
    return check;
}

void or_cache_load_all(or_cache *c)
{
    uint64_t total = 0;
    for (uint32_t l = 0; l < c->cfg.nb_levels; ++l)
        total += (uint64_t)c->g.level_dim[l][0] * c->g.level_dim[l][1] * c->g.level_dim[l][2];
    uint32_t *asked = (uint32_t *)malloc(total * 4 * sizeof(uint32_t));
    uint64_t n = 0;
    const uint32_t X = c->g.mrpd_dim[0];
    for (uint32_t z = 0; z < c->g.mrpd_dim[2]; ++z)
        for (uint32_t y = 0; y < c->g.mrpd_dim[1]; ++y)
            for (uint32_t x = 0; x < X; ++x) {
                uint32_t level = 0;
                while (x >= c->g.level_coord[level] + c->g.level_dim[level][0]) ++level;
                if (y >= c->g.level_dim[level][1] || z >= c->g.level_dim[level][2]) continue;
                asked[4 * n + 0] = x - c->g.level_coord[level];
                asked[4 * n + 1] = y;
                asked[4 * n + 2] = z;
                asked[4 * n + 3] = level;
                ++n;
            }
    handle_request(c, asked, (uint32_t)n);
    load_new_bricks(c);
    free(asked);
}

float or_completude(const or_cache *c)
{
    /* cache_use (CacheManagerKernels.cu:117-125) */
    uint32_t k = 0;
    for (uint32_t i = 0; i < c->g.nslots; ++i)
        if (c->lru[i].vol_pos[0] != -1.f) ++k;
    return (float)k / (float)c->g.nslots;
}

/* ===================================================================== DVR */

void or_view_tables(or_view *v, const or_cfg *cfg)
{
    /* VolumeRayCaster.cu:136-161 */
    const float bmin[3] = {-1.f, -1.f, -1.f}, bmax[3] = {1.f, 1.f, 1.f};
    float size[3];
    for (int a = 0; a < 3; ++a) size[a] = bmax[a] - bmin[a];
    const float smin = fminf(size[0], fminf(size[1], size[2]));
    v->tstep = smin / 256.f;
    v->nb_lods = cfg->nb_levels;
    for (uint32_t i = 0; i < cfg->nb_levels; ++i) {
        for (int a = 0; a < 3; ++a) v->inv_levels_size[i][a] = smin / (float)cfg->initial[i][a];
        const uint32_t m = MIN(cfg->initial[i][0], MIN(cfg->initial[i][1], cfg->initial[i][2]));
        v->lod_step[i] = (smin / (float)m) / 2.f;
    }
}

void or_tf_samples(const float *knots, uint32_t nk, uint32_t n, float *out)
{
    /* TransferFunction.cpp:121-157; lerp(a,b,t) = a + (b-a)*t, invLerp(a,b,v)
     * = (v-a)/(b-a), all in float. */
    for (uint32_t i = 0; i < n; ++i) {
        const float x = (float)i / (float)n;
        float *o = out + 4 * i;
        if (nk == 0) { o[0] = o[1] = o[2] = o[3] = 0.f; continue; }
        uint32_t pos = 0;
        while (pos < nk && !(knots[5 * pos] > x)) ++pos;
        float prev[5], next[5];
        if (pos == 0) {
            prev[0] = 0.f; prev[1] = 0.f;
            memcpy(prev + 2, knots + 5 * pos + 2, 3 * sizeof(float));
        } else {
            memcpy(prev, knots + 5 * (pos - 1), 5 * sizeof(float));
        }
        if (pos == nk) {
            next[0] = 1.f; next[1] = 0.f;
            memcpy(next + 2, knots + 5 * (nk - 1) + 2, 3 * sizeof(float));
        } else {
            memcpy(next, knots + 5 * pos, 5 * sizeof(float));
        }
        const float t = (x - prev[0]) / (next[0] - prev[0]);
        o[0] = prev[2] + (next[2] - prev[2]) * t;
        o[1] = prev[3] + (next[3] - prev[3]) * t;
        o[2] = prev[4] + (next[4] - prev[4]) * t;
        o[3] = prev[1] + (next[1] - prev[1]) * t;
    }
}

typedef struct { float x, y, z; } f3;

static inline f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 muls(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross3(f3 a, f3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* normalize = v * (1/sqrt(dot(v,v))), correctly rounded sqrt and divide.
 * (The reference's helper_math.h uses an approximate reciprocal square root
 * whose exact rounding is vendor-specific; this build fixes one IEEE rule for
 * both the device and this oracle.) */
static inline f3 norm3(f3 v) { const float inv = 1.0f / sqrtf(dot3(v, v)); return muls(v, inv); }
static inline f3 mulm(const float *m, f3 v)
{
    return mk(v.x * m[0] + v.y * m[1] + v.z * m[2],
              v.x * m[4] + v.y * m[5] + v.z * m[6],
              v.x * m[8] + v.y * m[9] + v.z * m[10]);
}

uint32_t or_compute_lod(const or_view *V, const float dleft[3], const float dup[3], float pdx, float pdy)
{
    /* compute_LOD (RayCasters.cu:551-567): level k is kept as soon as the
     * next level's voxel spacing along either screen axis reaches the pixel
     * footprint on that axis. */
    const f3 l = mk(dleft[0], dleft[1], dleft[2]), u = mk(dup[0], dup[1], dup[2]);
    int32_t k;
    for (k = 0; (uint32_t)k < V->nb_lods - 1; ++k) {
        const float *ns = V->inv_levels_size[k + 1];
        const f3 a = mul(l, mk(ns[0], ns[1], ns[2]));
        const f3 b = mul(u, mk(ns[0], ns[1], ns[2]));
        const float sl = sqrtf(dot3(a, a));
        const float su = sqrtf(dot3(b, b));
        if (sl >= pdx * 1.0f || su >= pdy * 1.0f) break;
    }
    return (uint32_t)k;
}

typedef struct {
    or_cache *c;
    const or_view *v;
    float *tbuf;
    double *cstate;
    float *rpos;
    uint32_t *pixels;
    int precision;
    int mop;                      /* 1: maximum projection (RayCastMOPImpl) */
    uint32_t next_row, end_row;
    pthread_mutex_t lock;
    uint64_t stats[4];
} dvr_job;

static inline double sample_box(const or_cache *c, uint32_t ex, uint32_t ey, uint32_t ez, float *fsum_out)
{
    /* tex3D, linear filter, normalised coordinates at e/size: per axis the
     * texels e-1 and e with weight 1/2 each (x, y clamp; z wraps because only
     * addressMode[0..1] are set, Texture3dDevice.hpp:104-122).  Returns the
     * exact sum of the 8 texels. */
    const uint32_t PX = c->g.pool_dim[0], PY = c->g.pool_dim[1], PZ = c->g.pool_dim[2];
    const uint32_t xs[2] = {ex == 0 ? 0 : ex - 1, ex};
    const uint32_t ys[2] = {ey == 0 ? 0 : ey - 1, ey};
    const uint32_t zs[2] = {ez == 0 ? PZ - 1 : ez - 1, ez};
    double s = 0.0;
    float fs = 0.f;
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i) {
                const uint64_t p = xs[i] + (uint64_t)PX * (ys[j] + (uint64_t)PY * zs[k]);
                if (c->cfg.bytes == 1) s += c->pool[p];
                else if (c->cfg.bytes == 2) s += ((const uint16_t *)c->pool)[p];
                else { const float f = ((const float *)c->pool)[p]; s += f; fs += f; }
            }
    if (fsum_out) *fsum_out = fs;
    return s;
}

int or_sample(const or_cache *c, uint32_t lod, const float np[3], double *sum8,
              uint32_t brick[3], uint32_t voxel[3])
{
    const or_geom *g = &c->g;
    const uint32_t inner = c->cfg.brick_size - 2 * c->cfg.covering;
    float pv[3];
    uint32_t b[3], v[3];
    for (int a = 0; a < 3; ++a) {
        const float pr = np[a] * g->ior[lod][a];
        pv[a] = pr * g->rdv[lod][a];
        /* the brick the walk addresses: the MRPD's, or the last table's */
        const uint32_t nb = g->nb_tables ? g->table[g->nb_tables - 1].virt[lod][a] : g->level_dim[lod][a];
        b[a] = (uint32_t)(pv[a] * (float)nb);
        v[a] = (uint32_t)(pv[a] * (float)g->virt_vox[lod][a]);
    }
    if (brick) memcpy(brick, b, sizeof(b));
    if (voxel) memcpy(voxel, v, sizeof(v));
    uint32_t e[4];
    walk((or_cache *)c, lod, pv, 0, 0, e);
    if (e[3] == 0) return 1;
    if (e[3] == 2) return 2;
    const uint32_t cov = c->cfg.covering;
    if (sum8) *sum8 = sample_box(c, cov + e[0] + v[0] % inner, cov + e[1] + v[1] % inner,
                                 cov + e[2] + v[2] % inner, NULL);
    return 0;
}

int or_probe(or_cache *c, int mode, uint32_t level, const float pos[3], double *raw)
{
    /* mode 0: K_CacheManager::get_normalized (K_CacheManager.hpp:267-361);
     * mode 1: K_CacheManager::get (:164-262), which on an Unmapped entry
     * raises the request only for the asked level and tries the next coarser
     * levels.  Mapped stamps the slot's usage (compute_element_position,
     * K_CacheDevice.hpp:316-334); raw is the 2x2x2 texel sum (mode 0) or the
     * texel at the element position (mode 1). */
    const or_geom *g = &c->g;
    const uint32_t asked = level, last = c->cfg.nb_levels - 1;
    const uint32_t bs = c->cfg.brick_size, cov = c->cfg.covering, inner = bs - 2 * cov;
    for (;;) {
        float real[3], virt[3];
        for (int a = 0; a < 3; ++a) {
            real[a] = pos[a] * g->ior[level][a];
            virt[a] = real[a] * g->rdv[level][a];
        }
        /* an Unmapped entry at any level of the walk is Unmapped (the
         * reference's get falls back to the coarser level after an Unmapped
         * MRPD entry, :241-252; after an Unmapped table entry it repeats the
         * lookup for ever, see walk) */
        uint32_t e[4];
        walk(c, level, virt, 1, c->ts, e);
        if (e[3] == 1) {
            const uint64_t ui = e[0] / bs + (uint64_t)c->cfg.cache_size[0] * (e[1] / bs + (uint64_t)c->cfg.cache_size[1] * (e[2] / bs));
            __atomic_store_n(&c->usage[ui], c->ts, __ATOMIC_RELAXED);
            uint32_t el[3];
            for (int a = 0; a < 3; ++a) el[a] = cov + e[a] + ((uint32_t)(virt[a] * (float)g->virt_vox[level][a])) % inner;
            if (mode == 0) {
                *raw = sample_box(c, el[0], el[1], el[2], NULL);
            } else {
                const uint64_t p = el[0] + (uint64_t)g->pool_dim[0] * (el[1] + (uint64_t)g->pool_dim[1] * el[2]);
                if (c->cfg.bytes == 1) *raw = c->pool[p];
                else if (c->cfg.bytes == 2) *raw = ((const uint16_t *)c->pool)[p];
                else *raw = ((const float *)c->pool)[p];
            }
            return 0;
        }
        if (e[3] == 2) return 2;
        if (mode == 0 || level == asked)
            __atomic_store_n(&c->req[or_mrpd_index(g, level, real)], c->ts, __ATOMIC_RELAXED);
        if (mode == 1 && level < last) {
            ++level;
            continue;
        }
        *raw = 0.0;
        return 1;
    }
}

static inline uint32_t pack_abgr(double r, double g, double b, double a)
{
    /* rgbaFloatToInt (RayCasters.cu:528-536) */
    const double s[4] = {r, g, b, a};
    uint32_t u[4];
    for (int i = 0; i < 4; ++i) {
        const double v = s[i] < 0.0 ? 0.0 : (s[i] > 1.0 ? 1.0 : s[i]);
        u[i] = (uint32_t)(v * 255.0);
    }
    return (u[3] << 24) | (u[2] << 16) | (u[1] << 8) | u[0];
}

static inline uint32_t pack_abgr_f(float r, float g, float b, float a)
{
    const float s[4] = {r, g, b, a};
    uint32_t u[4];
    for (int i = 0; i < 4; ++i) {
        const float v = fminf(fmaxf(s[i], 0.f), 1.f);
        u[i] = (uint32_t)(v * 255.f);
    }
    return (u[3] << 24) | (u[2] << 16) | (u[1] << 8) | u[0];
}

static void dvr_pixel(dvr_job *job, uint32_t x, uint32_t y, uint64_t st[4])
{
    or_cache *c = job->c;
    const or_view *V = job->v;
    const or_geom *g = &c->g;
    const uint32_t W = V->width, H = V->height;
    const uint32_t ts = c->ts;
    const uint32_t bs = c->cfg.brick_size, cov = c->cfg.covering;
    const uint32_t inner = bs - 2 * cov;
    const float *M = V->inv_mv;

    /* RayCasters.cu:57-84 */
    uint32_t xn = x + 1;
    if (xn > W) xn = x - 1;
    uint32_t yn = y + 1;
    if (yn > H) yn = y - 1;
    const float u = (((float)x + 0.5f) / (float)W) * 2.0f - 1.0f;
    const float un = (((float)xn + 0.5f) / (float)W) * 2.0f - 1.0f;
    const float vv = (((float)y + 0.5f) / (float)H) * 2.0f - 1.0f;
    const float vn = (((float)yn + 0.5f) / (float)H) * 2.0f - 1.0f;
    const f3 origin = mk(M[3], M[7], M[11]);
    const f3 dir = mulm(M, norm3(mk(u, vv, -2.f)));
    const f3 dxn = mulm(M, norm3(mk(un, vv, -2.f)));
    const f3 dyn = mulm(M, norm3(mk(u, vn, -2.f)));
    const f3 dleft = norm3(cross3(dir, dyn));
    const f3 dup = norm3(cross3(dir, dxn));

    /* intersectBox (RayCasters.cu:506-525) */
    const f3 invr = mk(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
    const f3 tbot = mul(invr, sub(mk(V->bbox_min[0], V->bbox_min[1], V->bbox_min[2]), origin));
    const f3 ttop = mul(invr, sub(mk(V->bbox_max[0], V->bbox_max[1], V->bbox_max[2]), origin));
    const f3 tmin = mk(fminf(ttop.x, tbot.x), fminf(ttop.y, tbot.y), fminf(ttop.z, tbot.z));
    const f3 tmax = mk(fmaxf(ttop.x, tbot.x), fmaxf(ttop.y, tbot.y), fmaxf(ttop.z, tbot.z));
    float tnear = fmaxf(fmaxf(tmin.x, tmin.y), fmaxf(tmin.x, tmin.z));
    const float tfar = fminf(fminf(tmax.x, tmax.y), fminf(tmax.x, tmax.z));
    const int hit = tfar > tnear;

    const uint64_t i = (uint64_t)y * W + x;
    const float tb = job->tbuf[i];
    tnear = fmaxf(tnear, tb);
    int unmapped = 0;
    double fc[4];
    float ff[4];
    memcpy(fc, job->cstate + 4 * i, sizeof(fc));
    for (int k = 0; k < 4; ++k) ff[k] = (float)fc[k];
    const int p64 = job->precision == 64;
    st[2] += 1;

    if (!hit) {
        fc[0] = fc[1] = fc[2] = 0.0;
        ff[0] = ff[1] = ff[2] = 0.f;
    } else {
        if (tnear < 0.f) tnear = 0.f;
        float t = tnear;
        f3 pos = add(origin, muls(dir, tnear));
        /* exact continuation (resume = 1, no reference counterpart): a ray
         * stopped at a missing brick resumes at the stored sample position */
        if (job->rpos && tb > 0.f && tb == t) pos = mk(job->rpos[4 * i], job->rpos[4 * i + 1], job->rpos[4 * i + 2]);
        const float epx = fabsf(dot3(sub(dir, dxn), dleft) * tfar);
        const float epy = fabsf(dot3(sub(dir, dyn), dup) * tfar);
        uint64_t prev[1 + OR_MAX_TABLES] = {UINT64_MAX, UINT64_MAX, UINT64_MAX, UINT64_MAX};
        uint32_t entry[4] = {0, 0, 0, 0};
        const float astop = V->alpha_stop;
        const uint32_t N = V->tf_size;
        for (;;) {
            const int go = p64 ? (fc[3] < (double)astop) : (ff[3] < astop);
            if (!(go && t < tfar)) break;
            const float alpha = t / tfar;
            const float pdx = epx * alpha, pdy = epy * alpha;
            f3 tp = add(muls(pos, 0.495f), mk(0.495f, 0.495f, 0.495f));
            tp.z = 1.0f - tp.z;
            const f3 np = mk(fmaxf(0.f, fminf(tp.x, 0.99f)), fmaxf(0.f, fminf(tp.y, 0.99f)),
                             fmaxf(0.f, fminf(tp.z, 0.99f)));
            const float dl[3] = {dleft.x, dleft.y, dleft.z}, du[3] = {dup.x, dup.y, dup.z};
            const uint32_t lod = or_compute_lod(V, dl, du, pdx, pdy);
            /* K_CacheManager::get_normalized (K_CacheManager.hpp:267-361) */
            const f3 pr = mul(np, mk(g->ior[lod][0], g->ior[lod][1], g->ior[lod][2]));
            const f3 pv = mul(pr, mk(g->rdv[lod][0], g->rdv[lod][1], g->rdv[lod][2]));
            /* the MRPD entry, then the table caches' (one-entry cache of the
             * walk's key, K_CacheManager.hpp:297-303) */
            const float pva[3] = {pv.x, pv.y, pv.z};
            uint64_t key[1 + OR_MAX_TABLES];
            walk_key(g, lod, pva, key);
            if (memcmp(key, prev, sizeof(key)) != 0) {
                walk(c, lod, pva, 1, ts, entry);
                memcpy(prev, key, sizeof(key));
                st[1] += 1;
            }
            if (entry[3] == 0) {
                /* raise_brick_request (K_CacheManager.hpp:365-396) with the
                 * real-volume position. */
                const uint32_t rx = g->level_coord[lod] + (uint32_t)(pr.x * (float)g->level_dim[lod][0]);
                const uint32_t ry = (uint32_t)(pr.y * (float)g->level_dim[lod][1]);
                const uint32_t rz = (uint32_t)(pr.z * (float)g->level_dim[lod][2]);
                const uint64_t ri = rx + (uint64_t)g->mrpd_dim[0] * (ry + (uint64_t)g->mrpd_dim[1] * rz);
                __atomic_store_n(&c->req[ri], ts, __ATOMIC_RELAXED);
                /* RayCasters.cu:138-147 */
                job->tbuf[i] = t;
                if (p64) memcpy(job->cstate + 4 * i, fc, sizeof(fc));
                else for (int k = 0; k < 4; ++k) job->cstate[4 * i + k] = ff[k];
                if (job->rpos) {
                    job->rpos[4 * i] = pos.x; job->rpos[4 * i + 1] = pos.y; job->rpos[4 * i + 2] = pos.z;
                    job->rpos[4 * i + 3] = t;
                }
                unmapped = 1;
                st[3] += 1;
                break;
            }
            if (entry[3] == 2) {
                /* Empty (RayCasters.cu:154-159) */
                t += V->tstep;
                if (t > tfar) break;
                pos = add(pos, muls(dir, V->tstep));
                continue;
            }
            /* Mapped: K_CacheDevice::compute_element_position
             * (K_CacheDevice.hpp:316-334) and the box-filtered fetch. */
            {
                const uint32_t ux = entry[0] / bs, uy = entry[1] / bs, uz = entry[2] / bs;
                const uint64_t ui = ux + (uint64_t)c->cfg.cache_size[0] * (uy + (uint64_t)c->cfg.cache_size[1] * uz);
                __atomic_store_n(&c->usage[ui], ts, __ATOMIC_RELAXED);
            }
            const uint32_t ex = cov + entry[0] + ((uint32_t)(pv.x * (float)g->virt_vox[lod][0])) % inner;
            const uint32_t ey = cov + entry[1] + ((uint32_t)(pv.y * (float)g->virt_vox[lod][1])) % inner;
            const uint32_t ez = cov + entry[2] + ((uint32_t)(pv.z * (float)g->virt_vox[lod][2])) % inner;
            float fsum = 0.f;
            const double sum = sample_box(c, ex, ey, ez, &fsum);
            st[0] += 1;
            const float kexp = V->lod_step[lod] / V->tstep;
            if (p64) {
                double s;
                if (c->cfg.bytes == 1) s = sum / 2040.0;
                else if (c->cfg.bytes == 2) s = sum / 524280.0;
                else s = sum / 8.0;
                /* tex1D, linear, normalised, clamp (RayCasters.cu:163,612-628) */
                const double xf = s * (double)N - 0.5;
                const double fl = floor(xf);
                const double a = xf - fl;
                int64_t i0 = (int64_t)fl, i1 = i0 + 1;
                i0 = i0 < 0 ? 0 : (i0 > (int64_t)N - 1 ? (int64_t)N - 1 : i0);
                i1 = i1 < 0 ? 0 : (i1 > (int64_t)N - 1 ? (int64_t)N - 1 : i1);
                double col[4];
                for (int k = 0; k < 4; ++k)
                    col[k] = (1.0 - a) * (double)V->tf[4 * i0 + k] + a * (double)V->tf[4 * i1 + k];
                col[3] = 1.0 - pow(1.0 - col[3], (double)kexp);
                if (col[3] > (double)0.001f) {
                    col[0] *= col[3]; col[1] *= col[3]; col[2] *= col[3];
                    const double w = 1.0 - fc[3];
                    for (int k = 0; k < 4; ++k) fc[k] += col[k] * w;
                }
            } else {
                float s;
                if (c->cfg.bytes == 1) s = (float)sum / 2040.0f;
                else if (c->cfg.bytes == 2) s = (float)sum / 524280.0f;
                else s = fsum / 8.0f;
                const float xf = s * (float)N - 0.5f;
                const float fl = floorf(xf);
                const float a = xf - fl;
                int64_t i0 = (int64_t)fl, i1 = i0 + 1;
                i0 = i0 < 0 ? 0 : (i0 > (int64_t)N - 1 ? (int64_t)N - 1 : i0);
                i1 = i1 < 0 ? 0 : (i1 > (int64_t)N - 1 ? (int64_t)N - 1 : i1);
                float col[4];
                for (int k = 0; k < 4; ++k)
                    col[k] = (1.0f - a) * V->tf[4 * i0 + k] + a * V->tf[4 * i1 + k];
                col[3] = 1.f - powf(1.f - col[3], kexp);
                if (col[3] > 0.001f) {
                    col[0] *= col[3]; col[1] *= col[3]; col[2] *= col[3];
                    const float w = 1.0f - ff[3];
                    float tmp[4];
                    for (int k = 0; k < 4; ++k) tmp[k] = col[k] * w;
                    for (int k = 0; k < 4; ++k) ff[k] += tmp[k];
                }
            }
            t += V->lod_step[lod];
            pos = add(pos, muls(dir, V->lod_step[lod]));
        }
        if (!unmapped) {
            if (job->rpos) {
                job->rpos[4 * i] = pos.x; job->rpos[4 * i + 1] = pos.y; job->rpos[4 * i + 2] = pos.z;
                job->rpos[4 * i + 3] = t;
            }
            job->tbuf[i] = t;
            if (p64) memcpy(job->cstate + 4 * i, fc, sizeof(fc));
            else for (int k = 0; k < 4; ++k) job->cstate[4 * i + k] = ff[k];
        }
    }
    job->pixels[i] = p64 ? pack_abgr(fc[0], fc[1], fc[2], fc[3]) : pack_abgr_f(ff[0], ff[1], ff[2], ff[3]);
}

/* RayCastMOPImpl (RayCasters.cu:201-349): the ray set-up, level of detail
 * and stepping of the DVR march (level steps, tstep over Empty bricks, stop
 * at the first Unmapped brick), with no early termination; the ray keeps the
 * largest normalised sample (colorStateBuffer[i].x) and the pixel is the
 * transfer function at it, without opacity correction or blending. */
static void mop_pixel(dvr_job *job, uint32_t x, uint32_t y, uint64_t st[4])
{
    or_cache *c = job->c;
    const or_view *V = job->v;
    const or_geom *g = &c->g;
    const uint32_t W = V->width, H = V->height;
    const uint32_t ts = c->ts;
    const uint32_t bs = c->cfg.brick_size, cov = c->cfg.covering;
    const uint32_t inner = bs - 2 * cov;
    const float *M = V->inv_mv;
    const int p64 = job->precision == 64;

    uint32_t xn = x + 1;
    if (xn > W) xn = x - 1;
    uint32_t yn = y + 1;
    if (yn > H) yn = y - 1;
    const float u = (((float)x + 0.5f) / (float)W) * 2.0f - 1.0f;
    const float un = (((float)xn + 0.5f) / (float)W) * 2.0f - 1.0f;
    const float vv = (((float)y + 0.5f) / (float)H) * 2.0f - 1.0f;
    const float vn = (((float)yn + 0.5f) / (float)H) * 2.0f - 1.0f;
    const f3 origin = mk(M[3], M[7], M[11]);
    const f3 dir = mulm(M, norm3(mk(u, vv, -2.f)));
    const f3 dxn = mulm(M, norm3(mk(un, vv, -2.f)));
    const f3 dyn = mulm(M, norm3(mk(u, vn, -2.f)));
    const f3 dleft = norm3(cross3(dir, dyn));
    const f3 dup = norm3(cross3(dir, dxn));
    const f3 invr = mk(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
    const f3 tbot = mul(invr, sub(mk(V->bbox_min[0], V->bbox_min[1], V->bbox_min[2]), origin));
    const f3 ttop = mul(invr, sub(mk(V->bbox_max[0], V->bbox_max[1], V->bbox_max[2]), origin));
    const f3 tmin = mk(fminf(ttop.x, tbot.x), fminf(ttop.y, tbot.y), fminf(ttop.z, tbot.z));
    const f3 tmax = mk(fmaxf(ttop.x, tbot.x), fmaxf(ttop.y, tbot.y), fmaxf(ttop.z, tbot.z));
    float tnear = fmaxf(fmaxf(tmin.x, tmin.y), fmaxf(tmin.x, tmin.z));
    const float tfar = fminf(fminf(tmax.x, tmax.y), fminf(tmax.x, tmax.z));
    const int hit = tfar > tnear;

    const uint64_t i = (uint64_t)y * W + x;
    const float tb = job->tbuf[i];
    tnear = fmaxf(tnear, tb);
    st[2] += 1;
    if (!hit) {
        job->pixels[i] = 0u;
        return;
    }
    if (tnear < 0.f) tnear = 0.f;
    float t = tnear;
    f3 pos = add(origin, muls(dir, tnear));
    if (job->rpos && tb > 0.f && tb == t) pos = mk(job->rpos[4 * i], job->rpos[4 * i + 1], job->rpos[4 * i + 2]);
    const float epx = fabsf(dot3(sub(dir, dxn), dleft) * tfar);
    const float epy = fabsf(dot3(sub(dir, dyn), dup) * tfar);
    uint64_t prev[1 + OR_MAX_TABLES] = {UINT64_MAX, UINT64_MAX, UINT64_MAX, UINT64_MAX};
    uint32_t entry[4] = {0, 0, 0, 0};
    /* maxOpacity: FP64 the sample s = sum / D in double, FP32 in float */
    double mx = job->cstate[4 * i];
    float mxf = (float)mx;
    const double D = c->cfg.bytes == 1 ? 2040.0 : (c->cfg.bytes == 2 ? 524280.0 : 8.0);
    int unmapped = 0;
    while (t < tfar) {
        const float alpha = t / tfar;
        const float pdx = epx * alpha, pdy = epy * alpha;
        f3 tp = add(muls(pos, 0.495f), mk(0.495f, 0.495f, 0.495f));
        tp.z = 1.0f - tp.z;
        const f3 np = mk(fmaxf(0.f, fminf(tp.x, 0.99f)), fmaxf(0.f, fminf(tp.y, 0.99f)),
                         fmaxf(0.f, fminf(tp.z, 0.99f)));
        const float dl[3] = {dleft.x, dleft.y, dleft.z}, du[3] = {dup.x, dup.y, dup.z};
        const uint32_t lod = or_compute_lod(V, dl, du, pdx, pdy);
        const f3 pr = mul(np, mk(g->ior[lod][0], g->ior[lod][1], g->ior[lod][2]));
        const f3 pv = mul(pr, mk(g->rdv[lod][0], g->rdv[lod][1], g->rdv[lod][2]));
        const float pva[3] = {pv.x, pv.y, pv.z};
        uint64_t key[1 + OR_MAX_TABLES];
        walk_key(g, lod, pva, key);
        if (memcmp(key, prev, sizeof(key)) != 0) {
            walk(c, lod, pva, 1, ts, entry);
            memcpy(prev, key, sizeof(key));
            st[1] += 1;
        }
        if (entry[3] == 0) {
            const uint32_t rx = g->level_coord[lod] + (uint32_t)(pr.x * (float)g->level_dim[lod][0]);
            const uint32_t ry = (uint32_t)(pr.y * (float)g->level_dim[lod][1]);
            const uint32_t rz = (uint32_t)(pr.z * (float)g->level_dim[lod][2]);
            const uint64_t ri = rx + (uint64_t)g->mrpd_dim[0] * (ry + (uint64_t)g->mrpd_dim[1] * rz);
            __atomic_store_n(&c->req[ri], ts, __ATOMIC_RELAXED);
            unmapped = 1;
            st[3] += 1;
            break;
        }
        if (entry[3] == 2) {
            t += V->tstep;
            if (t > tfar) break;
            pos = add(pos, muls(dir, V->tstep));
            continue;
        }
        {
            const uint32_t ux = entry[0] / bs, uy = entry[1] / bs, uz = entry[2] / bs;
            const uint64_t ui = ux + (uint64_t)c->cfg.cache_size[0] * (uy + (uint64_t)c->cfg.cache_size[1] * uz);
            __atomic_store_n(&c->usage[ui], ts, __ATOMIC_RELAXED);
        }
        const uint32_t ex = cov + entry[0] + ((uint32_t)(pv.x * (float)g->virt_vox[lod][0])) % inner;
        const uint32_t ey = cov + entry[1] + ((uint32_t)(pv.y * (float)g->virt_vox[lod][1])) % inner;
        const uint32_t ez = cov + entry[2] + ((uint32_t)(pv.z * (float)g->virt_vox[lod][2])) % inner;
        float fsum = 0.f;
        const double sum = sample_box(c, ex, ey, ez, &fsum);
        st[0] += 1;
        /* maxOpacity = max(maxOpacity, sample) */
        const double s = sum / D;
        if (s > mx) mx = s;
        const float sf = c->cfg.bytes == 4 ? fsum / 8.0f : (float)sum / (float)D;
        if (sf > mxf) mxf = sf;
        t += V->lod_step[lod];
        pos = add(pos, muls(dir, V->lod_step[lod]));
    }
    (void)unmapped;
    job->tbuf[i] = t;
    job->cstate[4 * i] = p64 ? mx : (double)mxf;
    if (job->rpos) {
        job->rpos[4 * i] = pos.x; job->rpos[4 * i + 1] = pos.y; job->rpos[4 * i + 2] = pos.z;
        job->rpos[4 * i + 3] = t;
    }
    /* classification: tex1D(tfTex, maxOpacity), linear, normalised, clamp */
    const uint32_t N = V->tf_size;
    if (p64) {
        const double xf = mx * (double)N - 0.5;
        const double fl = floor(xf), a = xf - fl;
        int64_t i0 = (int64_t)fl, i1 = i0 + 1;
        i0 = i0 < 0 ? 0 : (i0 > (int64_t)N - 1 ? (int64_t)N - 1 : i0);
        i1 = i1 < 0 ? 0 : (i1 > (int64_t)N - 1 ? (int64_t)N - 1 : i1);
        double col[4];
        for (int k = 0; k < 4; ++k) col[k] = (1.0 - a) * (double)V->tf[4 * i0 + k] + a * (double)V->tf[4 * i1 + k];
        job->pixels[i] = pack_abgr(col[0], col[1], col[2], col[3]);
    } else {
        const float xf = mxf * (float)N - 0.5f;
        const float fl = floorf(xf), a = xf - fl;
        int64_t i0 = (int64_t)fl, i1 = i0 + 1;
        i0 = i0 < 0 ? 0 : (i0 > (int64_t)N - 1 ? (int64_t)N - 1 : i0);
        i1 = i1 < 0 ? 0 : (i1 > (int64_t)N - 1 ? (int64_t)N - 1 : i1);
        float col[4];
        for (int k = 0; k < 4; ++k) col[k] = (1.0f - a) * V->tf[4 * i0 + k] + a * V->tf[4 * i1 + k];
        job->pixels[i] = pack_abgr_f(col[0], col[1], col[2], col[3]);
    }
}

static void *dvr_worker(void *arg)
{
    dvr_job *job = (dvr_job *)arg;
    uint64_t st[4] = {0, 0, 0, 0};
    for (;;) {
        pthread_mutex_lock(&job->lock);
        const uint32_t y = job->next_row++;
        pthread_mutex_unlock(&job->lock);
        if (y >= job->end_row) break;
        for (uint32_t x = 0; x < job->v->width; ++x) {
            if (job->mop) mop_pixel(job, x, y, st);
            else dvr_pixel(job, x, y, st);
        }
    }
    pthread_mutex_lock(&job->lock);
    for (int k = 0; k < 4; ++k) job->stats[k] += st[k];
    pthread_mutex_unlock(&job->lock);
    return NULL;
}

void or_render_dvr(or_cache *c, const or_view *v, float *tbuf, double *cstate,
                   uint32_t *pixels, int precision, int nthreads, uint64_t stats[4])
{
    or_render_dvr_rows(c, v, tbuf, cstate, pixels, precision, nthreads, 0, v->height, stats);
}

void or_render_dvr_rows(or_cache *c, const or_view *v, float *tbuf, double *cstate,
                        uint32_t *pixels, int precision, int nthreads, uint32_t y0, uint32_t rows,
                        uint64_t stats[4])
{
    or_render_dvr_resume(c, v, tbuf, cstate, NULL, pixels, precision, nthreads, y0, rows, stats);
}

static void render_rows(or_cache *c, const or_view *v, float *tbuf, double *cstate, float *rpos,
                        uint32_t *pixels, int precision, int mop, int nthreads, uint32_t y0, uint32_t rows,
                        uint64_t stats[4])
{
    dvr_job job;
    memset(&job, 0, sizeof(job));
    job.c = c; job.v = v; job.tbuf = tbuf; job.cstate = cstate; job.rpos = rpos; job.pixels = pixels;
    job.precision = precision;
    job.mop = mop;
    job.next_row = y0;
    job.end_row = y0 + rows > v->height ? v->height : y0 + rows;
    pthread_mutex_init(&job.lock, NULL);
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int k = 1; k < nthreads; ++k) pthread_create(&th[k], NULL, dvr_worker, &job);
    dvr_worker(&job);
    for (int k = 1; k < nthreads; ++k) pthread_join(th[k], NULL);
    free(th);
    pthread_mutex_destroy(&job.lock);
    if (stats) memcpy(stats, job.stats, sizeof(job.stats));
}

void or_render_dvr_resume(or_cache *c, const or_view *v, float *tbuf, double *cstate, float *rpos,
                          uint32_t *pixels, int precision, int nthreads, uint32_t y0, uint32_t rows,
                          uint64_t stats[4])
{
    render_rows(c, v, tbuf, cstate, rpos, pixels, precision, 0, nthreads, y0, rows, stats);
}

void or_render_mop(or_cache *c, const or_view *v, float *tbuf, double *cstate, float *rpos, uint32_t *pixels,
                   int precision, int nthreads, uint64_t stats[4])
{
    render_rows(c, v, tbuf, cstate, rpos, pixels, precision, 1, nthreads, 0, v->height, stats);
}
