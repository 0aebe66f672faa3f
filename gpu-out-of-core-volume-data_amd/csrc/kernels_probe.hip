// kernels_probe.hip -- gc_cache_probe: the device page-table API of
// include/gc_cache_device.h (gc_get_normalized, gc_get) run over a list of
// (level, position) probes, one thread per probe, each with its own copy of
// the kernel object (the reference's per-thread K_CacheManager,
// K_CacheManager.hpp:164-361).  For tests and tools: it exercises exactly
// the header user kernels compile against.
#include "gc_internal.hpp"
#include "gc_cache_device.h"

namespace gc {
namespace {

template <typename T>
__global__ void k_probe(gc_kcache k, uint32_t mode, const float *pos, const uint32_t *levels, uint32_t n,
                        float *values, uint32_t *status)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float3 p = make_float3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    if (mode == 0) {
        float v = values[i];
        status[i] = (uint32_t)gc_get_normalized<T>(k, levels[i], p, v);
        values[i] = v;
    } else {
        T v = T(0);
        const gc_voxel_status st = gc_get<T>(k, levels[i], p, v);
        status[i] = (uint32_t)st;
        if (st != GC_VOXEL_EMPTY) values[i] = (float)v;
    }
}

} // namespace
} // namespace gc

using namespace gc;

extern "C" int gc_cache_probe(gc_cache *c, uint32_t mode, const float *d_pos, const uint32_t *d_levels, uint32_t n,
                              float *d_values, uint32_t *d_status)
{
    if (!c || !d_pos || !d_levels || !d_values || !d_status) return fail(GC_ERR_ARG, "null argument");
    if (mode > 1) return fail(GC_ERR_ARG, "mode must be 0 (get_normalized) or 1 (get)");
    if (n == 0) return GC_OK;
    gc_cache_view view{};
    GC_TRY(gc_cache_view_get(c, &view));
    gc_geometry g{};
    GC_TRY(gc_cache_geometry(c, &g));
    void *stream = nullptr;
    GC_TRY(gc_cache_stream(c, &stream));
    const gc_kcache k = gc_kcache_from_view(&view);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid((n + 255) / 256), block(256);
    switch (g.encoded_bytes) {
    case 1: k_probe<uint8_t><<<grid, block, 0, s>>>(k, mode, d_pos, d_levels, n, d_values, d_status); break;
    case 2: k_probe<uint16_t><<<grid, block, 0, s>>>(k, mode, d_pos, d_levels, n, d_values, d_status); break;
    default: k_probe<float><<<grid, block, 0, s>>>(k, mode, d_pos, d_levels, n, d_values, d_status); break;
    }
    GC_HIP(hipGetLastError());
    GC_HIP(hipStreamSynchronize(s));
    return GC_OK;
}
