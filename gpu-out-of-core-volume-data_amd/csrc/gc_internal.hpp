// gc_internal.hpp -- definitions shared by the HIP kernels and the host
// runtime of libgcvol.so.  Everything here is internal; the public surface is
// include/gcvol.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "gcvol.h"

namespace gc {

// ------------------------------------------------------------------ errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define GC_HIP(expr)                                                                \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                       \
            return ::gc::fail(GC_ERR_HIP, std::string(#expr) + ": " +              \
                                              hipGetErrorString(e_));               \
    } while (0)

// propagate a non-GC_OK status
#define GC_TRY(expr)                          \
    do {                                      \
        int rc_ = (expr);                     \
        if (rc_ != GC_OK) return rc_;         \
    } while (0)

// ------------------------------------------------- FP32 bookkeeping formulas
// Evaluated identically on host and device (the library is compiled with
// -ffp-contract=off and IEEE-rounded divide/sqrt).

// Page-table position of brick `b` at `level`: b / (virt * (1/(bs-2c))) + FLT_EPSILON
// (CacheManagerKernels.cu:75-86, epsilon from CacheManager.hpp:560).
__host__ __device__ inline void volume_position(const gc_geometry &g, uint32_t level,
                                                const uint32_t b[3], float out[3])
{
    for (int a = 0; a < 3; ++a) {
        const float den = (float)g.virt_vox[level][a] * g.one_over[a];
        const float q = (float)b[a] / den;
        out[a] = q + 0x1p-23f; // FLT_EPSILON
    }
}

// MRPD cell of a normalised position (K_MultiResolutionPageDirectory.hpp:168-180).
__host__ __device__ inline uint64_t mrpd_index(const gc_geometry &g, uint32_t level,
                                               const float p[3])
{
    const uint32_t x = g.level_coord[level] + (uint32_t)(p[0] * (float)g.level_dim[level][0]);
    const uint32_t y = (uint32_t)(p[1] * (float)g.level_dim[level][1]);
    const uint32_t z = (uint32_t)(p[2] * (float)g.level_dim[level][2]);
    return (uint64_t)x + (uint64_t)g.mrpd_dim[0] * ((uint64_t)y + (uint64_t)g.mrpd_dim[1] * z);
}

// Per-slot LRU metadata.  The LRU order itself is a permutation of slot ids;
// the slot's voxel origin in the reference's 3-D pool (its cachePosition) is
// a function of the slot id, so only level and volumePosition are stored.
struct SlotMeta {
    float vol[3];
    uint32_t level;
};

__host__ __device__ inline void slot_origin(const gc_geometry &g, uint32_t s, uint32_t o[3])
{
    const uint32_t cx = g.cache_size[0], cy = g.cache_size[1];
    o[0] = (s % cx) * g.brick_size;
    o[1] = ((s / cx) % cy) * g.brick_size;
    o[2] = (s / (cx * cy)) * g.brick_size;
}

// ------------------------------------------------------ kernel entry points
// (implemented in kernels.hip; all enqueue on `stream` and return hipError_t)

struct LruScratch {
    uint64_t *masks;        // one bit per LRU entry
    uint32_t *counts;       // per-tile used counts, then exclusive offsets
    uint32_t *total;        // [0] used total
    uint32_t nblocks;
};

hipError_t launch_lru_update(const uint32_t *order_in, uint32_t *order_out, const uint32_t *usage,
                             uint32_t n, uint32_t ts, const LruScratch &s, hipStream_t stream);

struct MissScratch {
    uint64_t *masks;
    uint32_t *counts;
    uint32_t *total;        // [0] number of requested cells (uncapped)
    uint32_t nblocks;
};

// Requested cells (req == ts) in linear order over [0, cells-1), first `cap`
// of them decoded to (x_in_level, y, z, level) into out4.
hipError_t launch_miss_compact(const uint32_t *req, uint64_t cells, uint32_t ts,
                               const gc_geometry *d_geom, uint32_t cap, uint32_t *out4,
                               const MissScratch &s, hipStream_t stream);

// The cell table is a compact copy of the page directory that the ray caster
// reads: one word per MRPD cell, slot << 2 | w (w = 0 unmapped, 1 mapped,
// 2 empty, as the entry's w).  Every kernel that writes the MRPD writes it.
__host__ __device__ inline uint32_t cell_word(uint32_t slot, uint32_t w) { return slot << 2 | w; }

hipError_t launch_deref(uint4 *mrpd, uint32_t *cells, const uint32_t *order, uint32_t nslots, uint32_t nne,
                        const SlotMeta *meta, const gc_geometry *d_geom, hipStream_t stream);

// Victims take the new level / position and receive the brick voxels from
// the device staging buffer (big bricks, Bricker_v2 layout).
hipError_t launch_assign_copy(const uint32_t *order, uint32_t nslots, uint32_t nne,
                              const uint32_t *asked4, const uint32_t *bb_index,
                              const uint32_t *bb_coord4, SlotMeta *meta, void *pool,
                              const void *staging, const gc_geometry *d_geom,
                              const uint32_t big_brick[3], uint32_t bytes, hipStream_t stream);

// Victims take the new level / position and receive bricks cut from device
// levels (fully-resident start-up).
hipError_t launch_assign_cut(const uint32_t *order, uint32_t nslots, uint32_t n,
                             const uint32_t *asked4, SlotMeta *meta, void *pool,
                             const void *const *d_levels, const uint32_t *d_level_dims,
                             const gc_geometry *d_geom, uint32_t bytes, hipStream_t stream);

// MRPD insert of the asked bricks (reference_new_bricks, MRPD only): the first
// nne map to their victim slot (w = 1), the rest are Empty (w = 2).  serial=1
// replays the reference's in-order single-thread insert.
hipError_t launch_insert(uint4 *mrpd, uint32_t *cells, const uint32_t *order, uint32_t nslots, uint32_t n,
                         uint32_t nne, const uint32_t *asked4, const gc_geometry *d_geom,
                         int serial, hipStream_t stream);

// ------------------------------------------------------ page-table caches
// The table caches between the MRPD and the pool (CacheManager.hpp:151-236).
// Table i (i = 0 next to the MRPD) holds blocks of block[i] entries; its
// entries are compact words, block-major (entry `loc` of block b at
// b * b3 + loc, loc x-fastest inside the block), child << 2 | w, the child
// being a block of table i + 1, or a brick slot for the last table.  The
// MRPD's cell table holds table-0 block ids.  Each block has a usage stamp,
// an LRU position (order: a permutation of block ids) and the level and
// position of the brick it was taken for (meta).
struct TableState {
    uint32_t *words[GC_MAX_TABLES];
    uint32_t *usage[GC_MAX_TABLES];
    uint32_t *order[GC_MAX_TABLES];
    SlotMeta *meta[GC_MAX_TABLES];
};

__host__ __device__ inline uint32_t word_state_of(uint32_t w) { return w & 3u; }
__host__ __device__ inline uint32_t table_b3(const gc_table_geometry &t) { return t.block[0] * t.block[1] * t.block[2]; }

// origin of block b of table t in its surface (the entry's x, y, z)
__host__ __device__ inline void block_origin(const gc_table_geometry &t, uint32_t b, uint32_t o[3])
{
    o[0] = (b % t.blocks[0]) * t.block[0];
    o[1] = ((b / t.blocks[0]) % t.blocks[1]) * t.block[1];
    o[2] = (b / (t.blocks[0] * t.blocks[1])) * t.block[2];
}

// MRPD cell of a position (K_MultiResolutionPageDirectory.hpp:168-180 over
// the MRPD's levels: the brick grid without table caches)
__host__ __device__ inline uint32_t top_index(const gc_geometry &g, uint32_t level, const float p[3])
{
    const uint32_t x = g.top_coord[level] + (uint32_t)(p[0] * (float)g.top_dim[level][0]);
    const uint32_t y = (uint32_t)(p[1] * (float)g.top_dim[level][1]);
    const uint32_t z = (uint32_t)(p[2] * (float)g.top_dim[level][2]);
    return x + g.top_cells_dim[0] * (y + g.top_cells_dim[1] * z);
}

// Entry of a position inside its table-t block: (uint)(p * virtualized) %
// BlockSize per axis (K_CacheDevice.hpp:221-233,316-334, covering 0), as the
// block-local index.
__host__ __device__ inline uint32_t table_local(const gc_table_geometry &t, uint32_t level, const float p[3])
{
    uint32_t l[3];
    for (int a = 0; a < 3; ++a) l[a] = ((uint32_t)(p[a] * (float)t.virt[level][a])) % t.block[a];
    return l[0] + t.block[0] * (l[1] + t.block[1] * l[2]);
}

// dereference_old_bricks through the tables: the entry that maps a victim
// slot is cleared if the walk still reaches it (see kernels_cache.hip).
hipError_t launch_deref_tables(uint32_t *cells, const TableState &ts_, const uint32_t *order, uint32_t nslots,
                               uint32_t nne, const SlotMeta *meta, const gc_geometry *d_geom, uint32_t ts,
                               hipStream_t stream);
// reference_new_bricks through the tables, one brick after the other (see
// kernels_cache.hip); unreferenced[0] counts the bricks left unreferenced
// because a table had no block left in this load.
hipError_t launch_reference_tables(uint4 *mrpd, uint32_t *cells, const TableState &ts_, const uint32_t *order,
                                   uint32_t nslots, uint32_t n, uint32_t nne, const uint32_t *asked4,
                                   const SlotMeta *meta, const gc_geometry *d_geom, uint32_t ts,
                                   uint32_t *unreferenced, hipStream_t stream);

// The same for one table cache, in parallel (see kernels_cache.hip): per
// brick k, cell[k] / loc[k] are its MRPD cell and entry inside its table
// block and rep[k] the first brick of the load in the same cell (computed on
// the host; no two bricks share a cell and an entry).  When the load would
// take as a victim the block of a cell it touches, the kernels leave the
// state alone and the serial referencing runs instead.
struct RefScratch {
    uint32_t *cell, *loc, *rep;          // cap: from the host
    uint32_t *w0, *rank, *ev;            // cap: MRPD words, victim ranks, event indices
    uint32_t *inv;                       // table 0: LRU position of each block
    uint32_t *flag;                      // [0] 1: the serial referencing runs
};
hipError_t launch_reference_table1(uint4 *mrpd, uint32_t *cells, const TableState &ts_, const uint32_t *order,
                                   uint32_t nslots, uint32_t n, uint32_t nne, const uint32_t *asked4,
                                   const SlotMeta *meta, const gc_geometry *d_geom, uint32_t ts,
                                   uint32_t *unreferenced, const RefScratch &rs, hipStream_t stream,
                                   uint32_t nblocks);

hipError_t launch_occupancy(const SlotMeta *meta, uint32_t nslots, uint32_t *count,
                            hipStream_t stream);

hipError_t launch_mip(const void *in, const uint32_t dims[3], uint32_t bytes, void *out,
                      hipStream_t stream);

// One level-l brick from its level-(l-1) children (mip on miss).
struct MipBrickArgs {
    int64_t g0[3];        // level-l voxel at brick index (0,0,0): b * inner - cov
    uint32_t dims[3];     // level-l extent (voxels outside are 0)
    uint32_t c0[3];       // child-brick index of children[0]
    uint32_t nc[3];       // child grid extent (children x-fastest)
    uint32_t bs, cov, inner;
};
hipError_t launch_mip_brick(const MipBrickArgs &a, uint32_t bytes, const void *children, void *out,
                            hipStream_t stream);

// Synthetic bricks on the device (k_synth_bricks): blob k's factor along axis
// a at level-l coordinate c is tab[axis[l][a].off + k * len + c - lo] for c in
// [lo, lo + len).
struct SynthAxis {
    uint64_t off;
    int64_t lo;
    uint32_t len, pad;
};
struct SynthLaunch {
    const float *tab;          // axis tables of every level
    const float *amp;          // (float) amplitude of every blob
    const uint32_t *keys;      // n x (bx, by, bz, level)
    uint8_t *dst;              // n bricks of bs^3 voxels
    uint32_t *below;           // n: 1 when every voxel is below thr
    uint32_t n, nb, bs, cov, bytes, seed, noise;
    uint32_t max_level;        // keys of coarser levels are skipped (left to another writer)
    float fmax, thr;
    SynthAxis axis[GC_MAX_LEVELS][3];
    uint32_t ext[GC_MAX_LEVELS][3];   // level extents (voxels outside are 0)
};
hipError_t launch_synth_bricks(const SynthLaunch &p, hipStream_t s);
// A box of voxels (x fastest) of one level of the synthetic volume, the
// values the bricks of k_synth_bricks hold at the same positions.
struct SynthBox {
    const float *tab, *amp;
    uint8_t *dst;
    int64_t o[3];              // first voxel (level coordinates)
    uint32_t n[3];             // extent
    uint32_t level, nb, bytes, seed, noise;
    float fmax;
    SynthAxis axis[3];
    uint32_t ext[3];
};
hipError_t launch_synth_box(const SynthBox &b, hipStream_t s);

// Bricks of keys (bx, by, bz, level) cut with the covering rule from
// resident device levels (levels[l], dims[l]; keys of levels below
// first_level skipped), written to out + u * bs^3 voxels, with below[u] =
// every voxel below thr.
struct CutKeys {
    const uint8_t *levels[GC_MAX_LEVELS];
    uint32_t dims[GC_MAX_LEVELS][3];
    const uint32_t *keys;
    uint8_t *out;
    uint32_t *below;
    uint32_t n, bs, cov, bytes, first_level;
    float thr;
};
hipError_t launch_cut_keys(const CutKeys &a, hipStream_t s);

hipError_t launch_brick_cut(const void *level, const uint32_t dims[3], uint32_t bytes,
                            uint32_t bs, uint32_t cov, const uint32_t *bricks_xyz, uint32_t n,
                            void *out, hipStream_t stream);

// ------------------------------------------------------------------ DVR
// A launch renders one or more views over one cache.  Work is split into
// packets of 64 rays: 64 consecutive entries of an explicit ray list
// (view << 26 | pixel), then the TWxTH pixel tiles of the starting views.
constexpr uint32_t DVR_MAX_VIEWS = 64;
constexpr uint32_t DVR_PIXEL_BITS = 26;
// Transfer functions in LDS: every distinct table of a launch gets a copy
// padded with one clamped entry on each side, when at most DVR_TF_LDS_TABLES
// tables of at most DVR_TF_LDS_MAX entries are used (u8/u16 voxels).
constexpr uint32_t DVR_TF_LDS_TABLES = 8;
constexpr uint32_t DVR_TF_LDS_MAX = 256;
constexpr uint32_t DVR_TF_OWNER = 0x80000000u;

struct DvrView {
    uint32_t width, height, tiles_x, packet_off;   // packet_off: first tile packet
    float m[12];                                   // d_invModelViewMatrix rows
    float bmin[3], bmax[3];
    uint32_t tf_lds_off;                           // entry offset of the LDS copy (| DVR_TF_OWNER: this
                                                   // view copies it in), when the launch has LDS tables
    uint32_t tf_size;
    uint32_t tile_pair;                            // 1: tiles of two packets (top, bottom half)
    float alpha_stop;
    const float4 *tf;
    uint32_t *pixels;
    float *tbuf;
    void *cstate;                                  // float4 (fp32) or double4 (fp64)
    float4 *rpos;                                  // exact continuation: per pixel the position (xyz) and
                                                   // t (w) of the sample a stopped ray resumes at; null:
                                                   // resume at pos = origin + dir * t (the reference)
    double *qbuf;                                  // FP32: per pixel the exact opacity product Q of the
                                                   // samples blended so far (read when tbuf > 0)
    double thr;                                    // FP32: a ray continues while Q > thr =
                                                   // (1 - alpha_stop)^(1/k0)
    uint8_t *stop;                                 // per pixel: 1 = the ray stopped at an Unmapped
                                                   // brick in the last launch (or null: not recorded)
    uint32_t stop_tile_off;                        // first 8x8 tile of this view in the stopped-ray
                                                   // collection after a launch
};

// The table caches as the ray casters walk them (K_CacheManager.hpp:297-339).
struct PageTables {
    uint32_t n;                                  // table caches; 0: the MRPD maps bricks
    uint32_t top_x, top_y;                       // MRPD cells per row, rows per plane
    uint32_t block[GC_MAX_TABLES][3];
    uint32_t b3[GC_MAX_TABLES];                  // entries per block
    uint32_t pow2[GC_MAX_TABLES];                // 1: every block dimension is a power of two
    const uint32_t *words[GC_MAX_TABLES];
    uint32_t *usage[GC_MAX_TABLES];
};

struct DvrLaunch {
    const uint4 *mrpd;
    const uint32_t *cells;      // cell table (cell_word per MRPD cell)
    PageTables pt;
    uint32_t *request;
    uint32_t *usage;
    const void *pool;
    const gc_geometry *geom;
    uint32_t ts;
    const uint32_t *ts_dev;     // timestamp read on the device instead (graph replay), or null
    float tstep;
    uint32_t nb_lods;
    float inv_levels[GC_MAX_LEVELS][3];
    float lod_step[GC_MAX_LEVELS];
    float kexp[GC_MAX_LEVELS];
    // exact opacity decisions of FP32 compositing (exact_alpha = 1): y =
    // x^(kexp[l] / kexp[0]) by esq[l] squarings (kexp[l] / kexp[0] =
    // 2^esq[l] on every level); a sample is skipped when y >= qskip =
    // (1 - 0.001)^(1/kexp[0])
    int32_t esq[GC_MAX_LEVELS];
    double qskip;
    uint32_t exact_alpha;
    DvrView view0;              // the view when `views` is null
    const DvrView *views;       // device table of n_views views, or null
    uint32_t n_views;
    uint32_t n_packets;         // tile packets, after the ray list's packets
    const uint32_t *list;       // explicit ray list, or null for all tiles
    const uint32_t *list_count; // device count of `list`
    uint32_t *queue;            // 8 per-XCD packet cursors, zero at launch
    gc_render_stats *stats;
    uint32_t refill_min;        // refill idle lanes once this many are idle
    uint32_t march;             // samples per lane between refill checks (1..8)
    uint32_t tf_lds_entries;    // float4 entries of the LDS transfer-function tables (0: read from HBM)
    uint32_t tf_lds_tables;     // distinct tables among them
    uint32_t brick_size;        // the geometry's brick size and covering, for the host's choice of kernel
    uint32_t covering;
    uint32_t exact_resume;      // 1: some view has rpos (exact continuation of stopped rays)
    uint32_t pipelined;         // 1: software-pipelined march (requires realDivVirtual == 1 on every
                                //    level and MRPD y*z cells, x cells below 2^24, and no table
                                //    caches; see launch_base)
    // Hand-over of long rays (pipelined lane march): a ray that has consumed
    // yield_after samples in the launch stops before the first sample of its
    // next mapped brick, stores its state as at a miss (without a request),
    // and is appended to yield_list (count *yield_count) with its exact
    // position at the same index of yield_pos (whatever the views' resume
    // mode), for a group march to continue before the cache update.  0: never.  With
    // yield_exited > 0 rays are handed over only once that many waves of the
    // launch have run out of rays (*exited counts them; the wave that
    // reaches the count raises the 64 flags go[32 k], one per 128-byte line,
    // which the waves poll), so the rays left in the last waves, the
    // launch's tail, go to the group march.
    uint32_t yield_after;
    uint32_t yield_exited;
    uint32_t *exited;
    uint32_t *go;
    uint32_t *yield_list;
    uint32_t *yield_count;
    float4 *yield_pos;
    uint32_t continuation;      // 1: `list` holds handed-over rays, their positions in yield_pos
                                //    (k_dvr_ray; not counted again as rays)
};

// Ray tiles: 0 8x8, 1 16x4, 2 32x2, 3 64x1 are one 64-ray packet each; 4
// 16x8 and 5 32x4 are two packets, the tile's top and bottom halves (16x4,
// 32x2), taken one after the other.  packet_kind: the packet shape (the ray
// caster's variant, 0..3); tile_pair: 1 for two-packet tiles.
constexpr uint32_t GC_TILES = 6;
inline uint32_t packet_kind(uint32_t tile) { return tile == 4 ? 1u : (tile == 5 ? 2u : tile & 3u); }
inline uint32_t tile_pair(uint32_t tile) { return tile >= 4 ? 1u : 0u; }
// packet shape (tw x th = 64)
inline void tile_shape(uint32_t tile, uint32_t &tw, uint32_t &th)
{
    static const uint32_t TW[4] = {8, 16, 32, 64};
    tw = TW[packet_kind(tile)];
    th = 64 / tw;
}
// tile columns of a view of width w, and its packets
inline uint32_t tile_cols(uint32_t tile, uint32_t w)
{
    uint32_t tw, th;
    tile_shape(tile, tw, th);
    return (w + tw - 1) / tw;
}
inline uint32_t tile_packets(uint32_t tile, uint32_t w, uint32_t h)
{
    uint32_t tw, th;
    tile_shape(tile, tw, th);
    const uint32_t rows = th << tile_pair(tile);
    return (tile_cols(tile, w) * ((h + rows - 1) / rows)) << tile_pair(tile);
}

// Dynamic LDS of a DVR launch (per-lane level-switch table and view table)
// and the persistent grid for it (workgroups of 4 waves).
uint32_t dvr_lds_bytes(uint32_t nb_lods, uint32_t n_views, uint32_t tf_entries);
uint32_t dvr_grid(uint32_t bytes, uint32_t precision, uint32_t tile, uint32_t n_packets, uint32_t lds_bytes,
                  uint32_t tf_tables);
// Group-per-ray march of a ray list (list launches only; see kernels_dvr.hip):
// persistent workgroups of 4 waves, each group of 16 or 64 lanes taking one
// ray at a time (the group size follows `rays`, the list length).
bool dvr_ray_supported(const DvrLaunch &p, uint32_t bytes);
uint32_t dvr_ray_lds_bytes(uint32_t n_views, uint32_t tf_entries);
// Assign LDS copies to the views' transfer functions (views[i].tf_lds_off);
// returns the table entries, or 0 when they do not fit (then tf_lds_off is
// unused and the kernels read the tables from HBM).
uint32_t dvr_assign_tf_lds(DvrView *views, uint32_t n, uint32_t bytes, uint32_t *tables);
// group-per-ray march: group = lanes per ray (16 or 64; 0 = by the list
// length: 64 up to 256 rays, else 16)
// Resumed-ray lists of up to RAY_G16_MAX, RAY_G8_MAX and RAY_G4_MAX rays run
// the group-per-ray march with 16, 8 and 4 lanes per ray (64 up to 256 rays);
// RAY_G4_MAX is the default longest list of the group march.  Captured
// batch graphs keep 16 lanes and RAY_G16_MAX.
constexpr uint32_t RAY_G16_MAX = 1u << 16, RAY_G8_MAX = 1u << 17, RAY_G4_MAX = 1u << 18;
uint32_t dvr_ray_group(uint32_t rays);
uint32_t dvr_ray_grid(uint32_t bytes, uint32_t precision, uint32_t rays, uint32_t lds_bytes, uint32_t group = 0);
hipError_t launch_dvr_ray(const DvrLaunch &p, uint32_t bytes, uint32_t precision, uint32_t rays, uint32_t grid,
                          hipStream_t stream, uint32_t group = 0);
hipError_t launch_dvr(const DvrLaunch &p, uint32_t bytes, uint32_t precision, uint32_t tile,
                      uint32_t grid, hipStream_t stream);
hipError_t launch_fill32(uint32_t *p, uint32_t n, uint32_t value, hipStream_t stream);
// Maximum projection over the launch's ray list (list_rays entries, host-known
// count) and the views' 8x8 tile packets (packet_off, tiles_x of tile 0).
hipError_t launch_mop(const DvrLaunch &p, uint32_t bytes, uint32_t precision, uint32_t list_rays, hipStream_t stream);

// The rays a launch stopped at an Unmapped brick (the views' stop flags),
// gathered into the next launch's ray list in screen order: 8x8 tile by 8x8
// tile over the views, each tile's rays contiguous, so that a 64-ray packet
// of the list holds neighbouring rays (coherent brick accesses), as a tile
// of the first launch does.  The flags are cleared as they are read.
struct CollectArgs {
    const DvrView *views;       // device table of n_views views, or null: `one`
    DvrView one;
    uint32_t n_views;
    uint32_t n_tiles;           // 8x8 tiles over the views (stop_tile_off ranges)
    uint32_t *out_list;         // entries view << DVR_PIXEL_BITS | pixel, appended
    uint32_t *out_count;
    uint32_t *view_stopped;     // per view: rays appended (or null)
};
inline uint32_t stop_tiles(uint32_t w, uint32_t h) { return ((w + 7) / 8) * ((h + 7) / 8); }
hipError_t launch_collect_stopped(const CollectArgs &a, hipStream_t stream);

// Exact opacity threshold of a view (FP32 compositing) and the check that a
// launch's thresholds are normal doubles (runtime.cpp).
double opacity_threshold(float alpha_stop, float k0);
void finish_exact(DvrLaunch &L, const DvrView *views, uint32_t n);
// The shared fields of a ray-caster launch over a cache (device state,
// timestamp, level tables), with the cache's stream, device and voxel bytes.
int cache_launch_base(gc_cache *c, DvrLaunch &L, hipStream_t &stream, int &device, uint32_t &bytes);

} // namespace gc
