// shim_demo -- the reference's application flow (Application.cu:58-128,
// VolumeRayCaster.cu:261-330) written against the C++ drop-in headers:
// load a .cfg, preprocess the raw volume, build a CacheManager<T>, sample it
// from a user kernel through K_CacheManager::get_normalized until every probe
// is resident, then render one frame with the plain single-view renderer.
//
//   shim_demo <volume.cfg> <out_prefix> [probes_per_axis] [table_block]
//
// table_block > 0 adds one page-table cache (CacheSize[1], BlockSize[1]) of
// cubic blocks of that many bricks per axis, with a block for every page
// block of every level.
//
// Writes <out_prefix>.samples (float per probe, level-major, x fastest),
// <out_prefix>.status (uint32 VoxelStatus per probe) and <out_prefix>.pixels
// (ABGR uint32, 128x128, FP64 compositing), and prints a summary line.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "tdns/gpucache/CacheManager.hpp"

using tdns::gpucache::CacheManager;
using tdns::gpucache::K_CacheManager;
using tdns::gpucache::VoxelStatus;

template <typename T>
__global__ void probe_kernel(K_CacheManager<T> cache, uint32_t level, uint32_t n, float *values, uint32_t *status)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * n * n) return;
    const float3 p = make_float3(((float)(i % n) + 0.5f) / (float)n, ((float)((i / n) % n) + 0.5f) / (float)n,
                                 ((float)(i / (n * n)) + 0.5f) / (float)n);
    float v = -1.f;
    status[i] = (uint32_t)cache.get_normalized(level, p, v);
    values[i] = v;
}

static void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(2);
    }
}

template <typename T>
static int run(const tdns::data::VolumeConfiguration &vc, tdns::data::MetaData &md, const std::string &out, uint32_t n,
               uint32_t table_block)
{
    // one brick slot per brick of every level: fully resident once loaded
    uint64_t bricks = 0;
    for (uint32_t l = 0; l < vc.NbLevels; ++l) {
        uint64_t nb = 1;
        for (int a = 0; a < 3; ++a) nb *= vc.RealVolumesSizes[l][a] / (vc.BrickSize[a] - 2 * vc.Covering[a]);
        bricks += nb;
    }
    const uint64_t brick_bytes = (uint64_t)vc.BrickSize[0] * vc.BrickSize[1] * vc.BrickSize[2] * vc.EncodedBytes;
    uint32_t shape[3];
    tdns::detail::check(gc_pool_shape(bricks * brick_bytes, brick_bytes, shape));
    tdns::data::CacheConfiguration cc;
    cc.CacheSize.emplace_back(shape[0], shape[1], shape[2]);
    cc.BlockSize.push_back(vc.BrickSize);
    if (table_block) {
        uint32_t blocks = 0;
        for (uint32_t l = 0; l < vc.NbLevels; ++l) {
            uint32_t nb = 1;
            for (int a = 0; a < 3; ++a) {
                const uint32_t inner = vc.BrickSize[a] - 2 * vc.Covering[a];
                const uint32_t b = (vc.RealVolumesSizes[l][a] + inner - 1) / inner;
                nb *= (b + table_block - 1) / table_block;
            }
            blocks += nb;
        }
        cc.CacheSize.emplace_back(blocks, 1u, 1u);
        cc.BlockSize.emplace_back(table_block, table_block, table_block);
    }
    cc.DataCacheFlags = 1;
    CacheManager<T> manager(vc, md, cc, 0);

    const uint32_t per_level = n * n * n;
    std::vector<float> values((size_t)per_level * vc.NbLevels);
    std::vector<uint32_t> status((size_t)per_level * vc.NbLevels);
    float *d_values = nullptr;
    uint32_t *d_status = nullptr;
    hip_check(hipMalloc(&d_values, per_level * sizeof(float)), "hipMalloc");
    hip_check(hipMalloc(&d_status, per_level * sizeof(uint32_t)), "hipMalloc");
    uint32_t passes = 0;
    for (uint32_t l = 0; l < vc.NbLevels; ++l) {
        for (uint32_t it = 0; it < 1000; ++it) {
            ++passes;
            probe_kernel<T><<<(per_level + 255) / 256, 256>>>(manager.to_kernel_object(), l, n, d_values, d_status);
            hip_check(hipGetLastError(), "probe_kernel");
            hip_check(hipDeviceSynchronize(), "probe_kernel");
            hip_check(hipMemcpy(&status[(size_t)l * per_level], d_status, per_level * 4, hipMemcpyDeviceToHost), "copy");
            hip_check(hipMemcpy(&values[(size_t)l * per_level], d_values, per_level * 4, hipMemcpyDeviceToHost), "copy");
            uint32_t unmapped = 0;
            for (uint32_t i = 0; i < per_level; ++i) unmapped += status[(size_t)l * per_level + i] == VoxelStatus::Unmapped;
            manager.update();
            tdns::detail::check(gc_cache_wait_idle(manager.handle()));
            if (unmapped == 0) break;
        }
    }
    hip_check(hipFree(d_values), "hipFree");
    hip_check(hipFree(d_status), "hipFree");

    // one frame of the plain renderer over the same cache
    const uint32_t W = 128, H = 128;
    gc_frame *frame = nullptr;
    tdns::detail::check(gc_frame_create(W, H, 64, &frame));
    float knots[6 * 5];
    tdns::detail::check(gc_tf_seeded(5, 6, knots));
    std::vector<float> tf(256 * 4);
    tdns::detail::check(gc_tf_samples(knots, 6, 256, tf.data()));
    void *d_tf = nullptr;
    tdns::detail::check(gc_malloc(&d_tf, tf.size() * sizeof(float)));
    tdns::detail::check(gc_memcpy_h2d(d_tf, tf.data(), tf.size() * sizeof(float)));
    gc_view view{};
    view.width = W;
    view.height = H;
    const float t[3] = {0.f, 0.f, -3.f};
    tdns::detail::check(gc_camera_matrix(20.f, 30.f, t, view.inv_mv));
    for (int a = 0; a < 3; ++a) {
        view.bbox_min[a] = -1.f;
        view.bbox_max[a] = 1.f;
    }
    view.alpha_stop = 0.95f;
    view.precision = 64;
    gc_frame_report rep{};
    tdns::detail::check(gc_render_frame(manager.handle(), &view, static_cast<const float *>(d_tf), 256, frame, 64, &rep));
    std::vector<uint32_t> pixels((size_t)W * H);
    tdns::detail::check(gc_frame_read_pixels(frame, pixels.data()));
    tdns::detail::check(gc_frame_destroy(frame));
    tdns::detail::check(gc_free(d_tf));

    std::vector<float> comp;
    manager.completude(comp);
    auto dump = [&](const std::string &name, const void *p, size_t bytes) {
        std::FILE *f = std::fopen((out + name).c_str(), "wb");
        if (!f || std::fwrite(p, 1, bytes, f) != bytes) {
            std::fprintf(stderr, "cannot write %s%s\n", out.c_str(), name.c_str());
            std::exit(2);
        }
        std::fclose(f);
    };
    dump(".samples", values.data(), values.size() * sizeof(float));
    dump(".status", status.data(), status.size() * sizeof(uint32_t));
    dump(".pixels", pixels.data(), pixels.size() * sizeof(uint32_t));
    std::printf("levels %u probes %u passes %u completude %.6f launches %u complete %u samples %llu\n", vc.NbLevels,
                per_level, passes, comp[0], rep.launches, rep.complete, (unsigned long long)rep.stats.mapped_samples);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <volume.cfg> <out_prefix> [probes_per_axis] [table_block]\n", argv[0]);
        return 1;
    }
    const uint32_t n = argc > 3 ? (uint32_t)std::atoi(argv[3]) : 16;
    const uint32_t table_block = argc > 4 ? (uint32_t)std::atoi(argv[4]) : 0;
    try {
        const tdns::data::VolumeConfiguration vc = tdns::data::load_volume_configuration(argv[1]);
        tdns::data::MetaData md;
        if (!md.load_raw_volume(vc, vc.VolumeDirectory + vc.VolumeFileName)) {
            std::fprintf(stderr, "cannot read %s%s\n", vc.VolumeDirectory.c_str(), vc.VolumeFileName.c_str());
            return 2;
        }
        switch (vc.EncodedBytes) {
        case 1: return run<uchar1>(vc, md, argv[2], n, table_block);
        case 2: return run<ushort1>(vc, md, argv[2], n, table_block);
        default: return run<float1>(vc, md, argv[2], n, table_block);
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
}
