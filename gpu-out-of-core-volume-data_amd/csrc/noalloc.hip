// noalloc.hip -- gc_render_views_noalloc: one ray-caster launch over up to
// NOALLOC_MAX_VIEWS views of one cache into caller-provided buffers, without
// any device allocation.
//
// The launch's scratch memory -- the persistent kernel's packet cursors and
// the device table of views -- lives in `scratch`, a file-scope device array
// of SCRATCH_BYTES that every GPU the library runs on holds until the process
// ends.  Its device address is looked up once per device (hipGetSymbolAddress,
// on the entry point's first call on that device) and only that address is
// handed to kernels.  Every call on a GPU shares that GPU's array: calls that
// overlap on one GPU (two host threads, or two streams) race on it.
//
// What the call writes is what gc_batch_begin + gc_batch_launch write for the
// same views (batch semantics: every view starts from an empty progressive
// state, one launch, no ray list); the view table is written into the scratch
// by small kernels (views passed by value), the packet cursors and the
// caller's t buffers and colour states are cleared by fill kernels, so a call
// records no memset or host-memory copy and can be captured in a hipGraph.
#include "gc_internal.hpp"

#include <mutex>

namespace gc {

namespace {

constexpr uint32_t NOALLOC_MAX_VIEWS = 64;
constexpr size_t CTL_BYTES = 256;   // [0..7] per-XCD packet cursors
constexpr size_t SCRATCH_BYTES = CTL_BYTES + (size_t)NOALLOC_MAX_VIEWS * sizeof(DvrView);

static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

// per-device address of `scratch` (null until looked up)
constexpr int MAX_DEVICES = 64;
std::mutex g_lock;
void *g_scratch[MAX_DEVICES] = {};

// The view table, a few views per kernel argument block.
constexpr uint32_t VIEWS_PER_PUT = 16;
struct ViewChunk {
    DvrView v[VIEWS_PER_PUT];
    uint32_t first, n;
};
__global__ void k_put_views(DvrView *dst, const ViewChunk c)
{
    const uint32_t i = threadIdx.x;
    if (i < c.n) dst[c.first + i] = c.v[i];
}

int scratch_of(int device, uint8_t **out)
{
    if (device < 0 || device >= MAX_DEVICES) return fail(GC_ERR_ARG, "device index outside 0..63");
    std::lock_guard<std::mutex> lk(g_lock);
    if (!g_scratch[device]) {
        // with `device` current: each GPU holds its own copy of the array
        void *p = nullptr;
        GC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(GC_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch[device] = p;
    }
    *out = static_cast<uint8_t *>(g_scratch[device]);
    return GC_OK;
}

} // namespace

} // namespace gc

using namespace gc;

extern "C" int gc_render_views_noalloc(gc_cache *c, const gc_view *views, const float *const *d_tfs,
                                       const uint32_t *tf_sizes, uint32_t n_views, uint32_t *const *d_pixels,
                                       float *const *d_tbufs, void *const *d_cstates, gc_render_stats *d_stats,
                                       void *stream)
{
    if (!c || !views || !d_tfs || !tf_sizes || !d_pixels || !d_tbufs || !d_cstates)
        return fail(GC_ERR_ARG, "null argument");
    if (n_views == 0) return fail(GC_ERR_ARG, "no view");
    if (n_views > NOALLOC_MAX_VIEWS)
        return fail(GC_ERR_ARG, "gc_render_views_noalloc: " + std::to_string(n_views) +
                                    " views exceed its scratch limit of " + std::to_string(NOALLOC_MAX_VIEWS));
    const uint32_t prec = views[0].precision;
    if (prec != 32 && prec != 64) return fail(GC_ERR_ARG, "precision must be 32 or 64");
    if (views[0].tile >= GC_TILES || views[0].march > 8) return fail(GC_ERR_ARG, "bad launch shape");
    DvrLaunch L{};
    hipStream_t cs = nullptr;
    int device = 0;
    uint32_t bytes = 0;
    GC_TRY(cache_launch_base(c, L, cs, device, bytes));
    GC_HIP(hipSetDevice(device));
    uint8_t *scr = nullptr;
    GC_TRY(scratch_of(device, &scr));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : cs;

    // the view table (batch-slot semantics, every view starting)
    DvrView h[NOALLOC_MAX_VIEWS];
    uint32_t packets = 0;
    for (uint32_t i = 0; i < n_views; ++i) {
        const gc_view &v = views[i];
        if (v.width == 0 || v.height == 0 || (uint64_t)v.width * v.height >= (1ull << DVR_PIXEL_BITS))
            return fail(GC_ERR_ARG, "view " + std::to_string(i) + ": bad size");
        if (v.precision != prec) return fail(GC_ERR_ARG, "view " + std::to_string(i) + ": precision differs");
        if (v.compositing != 0) return fail(GC_ERR_ARG, "view " + std::to_string(i) + ": DVR only (compositing 0)");
        if (!d_tfs[i] || tf_sizes[i] == 0 || !d_pixels[i] || !d_tbufs[i] || !d_cstates[i])
            return fail(GC_ERR_ARG, "view " + std::to_string(i) + ": missing buffer");
        DvrView &D = h[i];
        D = DvrView{};
        D.width = v.width;
        D.height = v.height;
        D.tiles_x = tile_cols(views[0].tile, v.width);
        D.tile_pair = tile_pair(views[0].tile);
        D.packet_off = packets;
        packets += tile_packets(views[0].tile, v.width, v.height);
        for (int k = 0; k < 12; ++k) D.m[k] = v.inv_mv[k];
        for (int k = 0; k < 3; ++k) {
            D.bmin[k] = v.bbox_min[k];
            D.bmax[k] = v.bbox_max[k];
        }
        D.alpha_stop = v.alpha_stop;
        D.tf_size = tf_sizes[i];
        D.tf = reinterpret_cast<const float4 *>(d_tfs[i]);
        D.pixels = d_pixels[i];
        D.tbuf = d_tbufs[i];
        D.cstate = d_cstates[i];
        D.rpos = nullptr;   // every ray starts in this launch
        D.qbuf = nullptr;
        D.thr = opacity_threshold(v.alpha_stop, L.kexp[0]);
    }
    uint32_t tf_tables = 0;
    L.tf_lds_entries = dvr_assign_tf_lds(h, n_views, bytes, &tf_tables);
    L.tf_lds_tables = tf_tables;
    finish_exact(L, h, n_views);

    // clear what starts at zero: packet cursors, the views' progressive state
    const uint32_t cwords = prec == 64 ? 8u : 4u;   // colour-state words per pixel
    GC_HIP(launch_fill32(reinterpret_cast<uint32_t *>(scr), 8, 0, s));
    for (uint32_t i = 0; i < n_views; ++i) {
        const uint32_t npix = h[i].width * h[i].height;
        GC_HIP(launch_fill32(reinterpret_cast<uint32_t *>(h[i].tbuf), npix, 0, s));
        GC_HIP(launch_fill32(static_cast<uint32_t *>(h[i].cstate), npix * cwords, 0, s));
    }
    DvrView *d_views = reinterpret_cast<DvrView *>(scr + CTL_BYTES);
    for (uint32_t first = 0; first < n_views; first += VIEWS_PER_PUT) {
        ViewChunk ch{};
        ch.first = first;
        ch.n = std::min(VIEWS_PER_PUT, n_views - first);
        for (uint32_t k = 0; k < ch.n; ++k) ch.v[k] = h[first + k];
        k_put_views<<<1, VIEWS_PER_PUT, 0, s>>>(d_views, ch);
        GC_HIP(hipGetLastError());
    }

    L.views = d_views;
    L.n_views = n_views;
    L.n_packets = packets;
    L.queue = reinterpret_cast<uint32_t *>(scr);
    L.stats = d_stats;
    L.refill_min = views[0].refill_min ? views[0].refill_min : 16;
    L.march = views[0].march ? views[0].march : 4;
    const uint32_t grid = dvr_grid(bytes, prec, views[0].tile, packets, dvr_lds_bytes(L.nb_lods, n_views, L.tf_lds_entries),
                                   tf_tables);
    GC_HIP(launch_dvr(L, bytes, prec, views[0].tile, grid, s));
    return GC_OK;
}
