/*
 * gcvol_oracle.h -- CPU restatement of GcCore's out-of-core volume path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the CPU
 * baseline; it is never linked into, loaded by, or called from the product
 * library (libgcvol.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.
 *
 * Every function cites the reference file:line whose rule it restates
 * (paths relative to the GcCore reference tree).  Pinning of each function
 * (closed forms, brute force, invariants) lives in tests/test_oracle_*.py.
 *
 * Conventions
 *   - 3-D arrays are x-fastest: idx = x + X*(y + Y*z)
 *     (K_DynamicArray3dDevice.hpp:119-123).
 *   - Voxel types: bytes = 1 (uint8), 2 (uint16), 4 (float32)
 *     (Application.cu:104-128).
 *   - FP32 bookkeeping formulas are evaluated in IEEE single precision with
 *     no contraction (the library is built with -ffp-contract=off).
 */
#ifndef GCVOL_ORACLE_H
#define GCVOL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_LEVELS 16
#define OR_MAX_TABLES 3

/* ---------------------------------------------------------------- geometry */

/* levelMax + 1 with levelMax = ceil(log(maxDim / bs) / log 2), integer
 * division maxDim/bs (Mipmapper.cu:441-443). */
uint32_t or_level_count(const uint32_t dims[3], uint32_t brick_size);

/* Level sizes: size_{l+1} = size_l >= r ? size_l / r : size_l
 * (Mipmapper.cu:583-597, fill_metaData). out: nb_levels x 3. */
void or_level_sizes(const uint32_t dims[3], const uint32_t ratio[3],
                    uint32_t nb_levels, uint32_t *out);

/* Bricker metadata (Bricker_v2.cpp:45-76): per level bricks per axis rounded
 * up to a multiple of the big-brick size, big bricks per axis and the real
 * (padded) level size.  Any output pointer may be NULL. */
void or_init_meta(uint32_t nb_levels, const uint32_t *levels, uint32_t brick_size,
                  uint32_t covering, const uint32_t big_brick[3],
                  uint32_t *nb_bricks, uint32_t *nb_big_bricks, uint32_t *real_sizes);

/* Pool shape from a byte budget (Application.cu:79-93). */
void or_pool_shape(uint64_t budget_bytes, uint64_t brick_bytes, uint32_t out[3]);

/* ------------------------------------------------------------ preprocessor */

/* Whole-level down-sampling, out[x,y,z] = mean of the r.x*r.y*r.z group,
 * floored for integer types (Mipmapper.cu:30-105): uint8 sums in uint64 and
 * divides; uint16 sums in double (start 0) and truncates; float sums in double
 * and rounds to float.  Output dims are floor(dims / r) per axis (the writer
 * drops the trailing odd voxel, Mipmapper.cu:44-45,248). */
void or_mip_reduce(const void *in, const uint32_t dims[3], uint32_t bytes,
                   const uint32_t ratio[3], void *out);

/* Cut one big brick (BB*bs per axis, small bricks at brickPos*bs) out of a
 * level with the covering rule of Bricker_v2.cpp:138-222: local voxel k of
 * small brick b maps to global b*(bs-2c) + k - c, zero outside the level. */
void or_brick_extract(const void *level, const uint32_t dims[3], uint32_t bytes,
                      uint32_t brick_size, uint32_t covering, const uint32_t big_brick[3],
                      const uint32_t big_brick_coord[3], void *out);
/* Bricks [b0, b1) of a level (x-fastest brick order, BigBrickSize 1), one
 * after the other: the CPU bricker's per-level loop (Bricker_v2.cpp:79-252). */
void or_brick_level(const void *level, const uint32_t dims[3], uint32_t bytes, uint32_t bs, uint32_t cov,
                    uint32_t b0, uint32_t b1, void *out);

/* ------------------------------------------------------------------- cache */

typedef struct {
    uint32_t brick_size;              /* BrickSize                          */
    uint32_t covering;                /* VoxelCovering                      */
    uint32_t bytes;                   /* NumberEncodedBytes                 */
    uint32_t big_brick[3];            /* BigBrickSizeX/Y/Z                  */
    uint32_t nb_levels;
    uint32_t initial[OR_MAX_LEVELS][3];   /* InitialVolumeSizes             */
    uint32_t real[OR_MAX_LEVELS][3];      /* RealVolumesSizes               */
    uint32_t cache_size[3];           /* CacheSize[0], in bricks            */
    uint32_t max_requests;            /* _nbMaxRequests (50)                */
    float    empty_threshold;         /* big brick Empty iff all voxels < t */
    /* Page-table caches between the MRPD and the data cache: CacheSize[k+1]
     * and BlockSize[k+1] of CacheConfiguration (CacheManager.hpp:138-141);
     * nb_tables = CacheSize.size() - 1, 0 for an MRPD that maps bricks. */
    uint32_t nb_tables;
    uint32_t table_cache[OR_MAX_TABLES][3];   /* blocks per axis            */
    uint32_t table_block[OR_MAX_TABLES][3];   /* entries per block per axis */
} or_cfg;

/* One page-table cache of the hierarchy (CacheDevice<uint4>,
 * CacheDevice.hpp:123-180): table[i] of or_geom is the reference's
 * _tableCaches[i] (CacheManager.hpp:223-233), i = 0 next to the MRPD, i.e.
 * hierarchy index nb_tables - i and configuration index nb_tables - 1 - i. */
typedef struct {
    uint32_t block[3];                      /* BlockSize                     */
    uint32_t blocks[3];                     /* CacheSize (blocks per axis)   */
    uint32_t surf[3];                       /* entries per axis: blocks*block */
    uint32_t nblocks;
    uint32_t virt[OR_MAX_LEVELS][3];        /* virtualizedNumberOfEntries    */
    uint32_t real[OR_MAX_LEVELS][3];        /* realNumberOfEntries           */
} or_table_geom;

typedef struct {
    uint32_t level_dim[OR_MAX_LEVELS][3];   /* realNumberOfEntries[1]        */
    uint32_t level_coord[OR_MAX_LEVELS];    /* MRPD x offset of each level   */
    uint32_t virt_vox[OR_MAX_LEVELS][3];    /* virtualizedNumberOfEntries[0] */
    float    ior[OR_MAX_LEVELS][3];         /* initialOverReal               */
    float    rdv[OR_MAX_LEVELS][3];         /* data cache real/virtual       */
    float    one_over[3];                   /* 1/(bs-2c)                     */
    uint32_t mrpd_dim[3];                   /* brick grid: the request buffer
                                               shape, and the MRPD's when
                                               nb_tables = 0                 */
    uint32_t pool_dim[3];                   /* voxels                        */
    uint32_t nslots;
    /* the MRPD (realNumberOfEntries.back(), MultiResolutionPageDirectory.hpp
     * :99-124): the brick grid above when nb_tables = 0 */
    uint32_t nb_tables;
    uint32_t top_dim[OR_MAX_LEVELS][3];
    uint32_t top_coord[OR_MAX_LEVELS];
    uint32_t top_cells_dim[3];
    or_table_geom table[OR_MAX_TABLES];
} or_geom;

/* LRU record (LruContent.hpp:10-15), 28 bytes. */
typedef struct {
    uint32_t cache_pos[3];
    uint32_t level;
    float    vol_pos[3];
} or_lru;

typedef struct or_cache or_cache;

/* CacheManager ctor geometry (CacheManager.hpp:157-285), any number of
 * table caches. */
void or_geometry(const or_cfg *cfg, or_geom *g);

/* Create a cache over a fully host-resident pyramid: levels[l] points to the
 * initial-size level-l voxels.  The pointers are borrowed. */
or_cache *or_cache_create(const or_cfg *cfg, const void *const *levels);
void      or_cache_destroy(or_cache *c);
const or_geom *or_cache_geom(const or_cache *c);

/* Raw state, exposed for comparison and for trace replay. */
uint32_t *or_cache_mrpd(or_cache *c);      /* 4 x uint32 per entry (top_cells_dim) */
uint32_t *or_cache_request(or_cache *c);   /* uint32 per brick cell (mrpd_dim) */
uint32_t *or_cache_usage(or_cache *c);     /* uint32 per slot, x-fastest    */
or_lru   *or_cache_lru(or_cache *c);       /* nslots records in LRU order   */
void     *or_cache_pool(or_cache *c);      /* pool_dim voxels, x-fastest    */
uint32_t  or_cache_timestamp(const or_cache *c);
/* Table cache i (i = 0 next to the MRPD): its uint4 entries (surf, x-fastest),
 * its per-block usage and its LRU (nblocks records, LRU order). */
uint32_t *or_cache_table(or_cache *c, uint32_t i);
uint32_t *or_cache_table_usage(or_cache *c, uint32_t i);
or_lru   *or_cache_table_lru(or_cache *c, uint32_t i);
/* Bricks of the last load that the table caches could not reference (their
 * tables had no victim block left in that load); 0 without table caches. */
uint32_t  or_cache_unreferenced(const or_cache *c);

/* CacheManager::update with a synchronous request handler
 * (CacheManager.hpp:319-354; the handler finishes inside notify). */
int  or_cache_update(or_cache *c);
/* The individual steps, for step-wise comparison. */
void or_lru_update(or_cache *c);                                        /* :387-406 */
uint32_t or_request_list(or_cache *c, uint32_t *out4, uint32_t cap);    /* :410-474 */
float or_completude(const or_cache *c);                                 /* :358-383 */
/* Last handled request: asked list (uint4, non-empty first) and counts. */
uint32_t or_asked(const or_cache *c, uint32_t *out4, uint32_t *nb_non_empty);

/* Map every brick as one load with the asked list = all bricks of every
 * level in MRPD linear order (no cap). */
void or_cache_load_all(or_cache *c);

/* FP32 page-table position of a brick (CacheManagerKernels.cu:75-86). */
void or_volume_position(const or_geom *g, uint32_t level, const uint32_t brick[3], float out[3]);
/* Brick-grid cell of a float position (K_MultiResolutionPageDirectory.hpp
 * :168-180 over realNumberOfEntries[1]): the MRPD cell when nb_tables = 0,
 * and the request-buffer cell (K_CacheManager.hpp:365-396). */
uint64_t or_mrpd_index(const or_geom *g, uint32_t level, const float pos[3]);
/* MRPD cell of a float position (the top of the hierarchy). */
uint64_t or_top_index(const or_geom *g, uint32_t level, const float pos[3]);

/* ------------------------------------------------------------------- DVR */

typedef struct {
    uint32_t width, height;
    float    inv_mv[12];               /* float4x3, rows (RayCasters_helpers.hpp:22-27) */
    float    bbox_min[3], bbox_max[3];
    float    tstep;                    /* marchingStep                   */
    float    alpha_stop;               /* 0.95f in the reference         */
    uint32_t nb_lods;
    float    inv_levels_size[OR_MAX_LEVELS][3];
    float    lod_step[OR_MAX_LEVELS];
    const float *tf;                   /* tf_size float4 (RGBA)          */
    uint32_t tf_size;
} or_view;

/* One page-table translation and box-filtered fetch at a normalised position
 * (K_CacheManager::get_normalized, K_CacheManager.hpp:267-361, without the
 * usage / request side effects).  Returns 0 Mapped, 1 Unmapped, 2 Empty;
 * sum8 receives the sum of the 8 texels (Mapped only); brick/voxel (may be
 * NULL) receive the FP32-derived brick index and in-level voxel index. */
int or_sample(const or_cache *c, uint32_t level, const float pos[3], double *sum8,
              uint32_t brick[3], uint32_t voxel[3]);

/* Host-side LOD tables (VolumeRayCaster.cu:136-161) for a [-1,1]^3 box. */
/* The device page-table API with its side effects (usage stamp, request at
 * the cache's timestamp): mode 0 get_normalized (K_CacheManager.hpp:267-361),
 * mode 1 get with the coarser-level fallback (:164-262).  Returns 0 Mapped,
 * 1 Unmapped, 2 Empty; raw: the 2x2x2 texel sum (mode 0) or the texel at the
 * element position (mode 1), 0 for Unmapped. */
int or_probe(or_cache *c, int mode, uint32_t level, const float pos[3], double *raw);
void or_view_tables(or_view *v, const or_cfg *cfg);
/* compute_LOD (RayCasters.cu:551-567): the level of a sample whose pixel
 * footprint is (pdx, pdy) along the screen axes dleft, dup. */
uint32_t or_compute_lod(const or_view *v, const float dleft[3], const float dup[3], float pdx, float pdy);

/* Transfer-function samples at x = i/N, piecewise linear between control
 * points with implicit end points (TransferFunction.cpp:121-157).
 * knots: nk x {value, alpha, r, g, b}, sorted by value. out: N float4. */
void or_tf_samples(const float *knots, uint32_t nk, uint32_t n, float *out);

/* One launch of RayCastDVRImpl (RayCasters.cu:36-198) over the cache.
 * Geometry (ray set-up, positions, LOD, page-table and voxel indices) is
 * evaluated in FP32 exactly as the reference orders it; value filtering,
 * transfer function, opacity correction and compositing are evaluated in
 * double when precision == 64, in float when precision == 32.
 * tbuf: W*H floats; cstate: W*H*4 doubles (colour state, RGBA).
 * pixels: W*H uint32 ABGR.  stats (may be NULL): {mapped samples, brick
 * visits, pixels, rays that stopped at an Unmapped brick}. */
void or_render_dvr(or_cache *c, const or_view *v, float *tbuf, double *cstate,
                   uint32_t *pixels, int precision, int nthreads, uint64_t stats[4]);
/* The same launch restricted to pixel rows [y0, y0+rows). */
void or_render_dvr_rows(or_cache *c, const or_view *v, float *tbuf, double *cstate,
                        uint32_t *pixels, int precision, int nthreads, uint32_t y0, uint32_t rows,
                        uint64_t stats[4]);
/* The same with exact continuation (the product's resume = 1 mode, which has
 * no reference counterpart): rpos (W*H*4 floats) receives the position and t
 * of the sample each ray stores its state at, and a ray whose t buffer holds
 * t > 0 resumes at that stored position instead of origin + dir * t.
 * rpos == NULL is or_render_dvr_rows. */
void or_render_dvr_resume(or_cache *c, const or_view *v, float *tbuf, double *cstate, float *rpos,
                          uint32_t *pixels, int precision, int nthreads, uint32_t y0, uint32_t rows,
                          uint64_t stats[4]);
/* RayCastMOPImpl (RayCasters.cu:201-349): maximum projection; cstate[4i] is
 * the normalised maximum (colorStateBuffer[i].x), the pixel TF(maximum). */
void or_render_mop(or_cache *c, const or_view *v, float *tbuf, double *cstate, float *rpos, uint32_t *pixels,
                   int precision, int nthreads, uint64_t stats[4]);

#ifdef __cplusplus
}
#endif
#endif
