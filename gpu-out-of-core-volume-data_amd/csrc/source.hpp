// source.hpp -- the brick-source interface behind gc_source (the role of
// BricksManager::get_brick, BricksManager.cpp:48-77): a source hands out big
// bricks -- BigBrickSize small bricks laid out as Bricker_v2.cpp:295-299 --
// by level and big-brick position.  Internal to libgcvol.so.
#pragma once

#include <memory>

#include "gc_internal.hpp"

namespace gc {

struct Source {
    gc_volume_cfg vc{};
    uint32_t bs = 0, cov = 0, bytes = 0, BB[3] = {1, 1, 1};
    uint64_t bb_vox = 0, bb_bytes = 0;
    virtual ~Source() = default;
    virtual int get(uint32_t level, const uint32_t bbc[3], void *dst) = 0;   // GC_BRICK_*
    virtual bool thread_safe() const { return true; }
    virtual uint32_t threads() const { return 1; }
    // Sources that write bricks on the device: n bricks (keys: bx, by, bz,
    // level) into d_dst on stream s, below[i] = every voxel of brick i below
    // thr; returns when the bricks are written.
    virtual bool on_device() const { return false; }
    virtual int get_device(const uint32_t *keys, uint32_t n, uint8_t *d_dst, float thr, uint8_t *below,
                           hipStream_t s)
    {
        (void)keys; (void)n; (void)d_dst; (void)thr; (void)below; (void)s;
        return fail(GC_ERR_STATE, "source does not write bricks on the device");
    }
    void init(const gc_volume_cfg &v)
    {
        vc = v;
        bs = v.brick_size[0];
        cov = v.covering[0];
        bytes = v.encoded_bytes;
        for (int a = 0; a < 3; ++a) BB[a] = v.big_brick_size[a];
        bb_vox = (uint64_t)bs * bs * bs * BB[0] * BB[1] * BB[2];
        bb_bytes = bb_vox * bytes;
    }
    // Offset (in voxels) of small brick (sx,sy,sz) inside a big brick
    // (get_iterator_from_indexes, Bricker_v2.cpp:295-299).
    uint64_t small_offset(uint32_t sx, uint32_t sy, uint32_t sz) const
    {
        const uint64_t BX = (uint64_t)bs * BB[0], BY = (uint64_t)bs * BB[1];
        return (uint64_t)sx * bs + BX * ((uint64_t)sy * bs + BY * (uint64_t)sz * bs);
    }
};

} // namespace gc

struct gc_source {
    std::unique_ptr<gc::Source> impl;
};
