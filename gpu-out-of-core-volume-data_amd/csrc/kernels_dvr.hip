// kernels_dvr.hip -- the DVR ray caster for gfx950 (RayCastDVRImpl,
// RayCasters.cu:36-198, with K_CacheManager::get_normalized,
// K_CacheManager.hpp:267-361, and K_CacheDevice::get_normalized,
// K_CacheDevice.hpp:191-201,316-334, inlined).
//
// Layout and mapping
//  * One lane per pixel; a workgroup covers a TWxTH ray tile (8x8, 16x8, 32x4
//    or 64x1), so neighbouring rays of a wave walk the same bricks.
//  * Workgroup ids are remapped so that consecutive tiles land on the same
//    XCD (workgroups are dealt round-robin over the 8 XCDs): each XCD walks a
//    contiguous band of the image and keeps its bricks in its own L2.
//  * Per-level tables and the transfer function live in LDS.
//  * The page entry of the current brick is cached per lane (the reference's
//    _previousCoords cache); the usage stamp is written once per brick visit.
//  * The pool is brick-major (slot * bs^3), so the 2x2x2 footprint of a
//    sample is at most four rows inside one brick.
//
// Arithmetic
//  * Ray set-up, positions, LOD choice and every page-table / voxel index are
//    FP32 in the reference's operation order (no contraction, IEEE divide and
//    square root), so the sample trace is reproducible on the host.
//  * The sample value is the exact mean of the 8 texels around the voxel
//    (the linear-filter texture fetch at an integer texel coordinate).
//  * Compositing is FP32 (precision 32, opacity correction with the hardware
//    exp2/log2) or FP64 (precision 64).
#include "gc_internal.hpp"

namespace gc {

namespace {

constexpr int TF_LDS = 256;

struct LevelLds {
    float ior[3], rdv[3], ldf[3], vvf[3], inv[3];
    uint32_t lcoord;
    float step, kexp;
};

struct F3 {
    float x, y, z;
};
__device__ inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline F3 operator*(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ inline F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline F3 cross(F3 a, F3 b)
{
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline F3 normalize(F3 v)
{
    const float inv = 1.0f / __fsqrt_rn(dot(v, v));
    return v * inv;
}
__device__ inline F3 mulm(const float *m, F3 v)
{
    return f3(v.x * m[0] + v.y * m[1] + v.z * m[2], v.x * m[4] + v.y * m[5] + v.z * m[6],
              v.x * m[8] + v.y * m[9] + v.z * m[10]);
}
__device__ inline float spacing(F3 d, const float *inv)
{
    const F3 a = d * f3(inv[0], inv[1], inv[2]);
    return __fsqrt_rn(dot(a, a));
}

// Sum of the 8 texels {e-1, e}^3 for a voxel strictly inside its brick
// (covering >= 1): q points at texel e.
template <int BYTES>
__device__ inline uint32_t box_sum_inner(const uint8_t *q, uint32_t bs)
{
    const uint32_t r = bs, s = bs * bs;
    if constexpr (BYTES == 1) {
        return (uint32_t)q[0] + q[-1] + q[-(int)r] + q[-(int)r - 1] + q[-(int)s] + q[-(int)s - 1] +
               q[-(int)s - (int)r] + q[-(int)s - (int)r - 1];
    } else {
        const uint16_t *p = reinterpret_cast<const uint16_t *>(q);
        return (uint32_t)p[0] + p[-1] + p[-(int)r] + p[-(int)r - 1] + p[-(int)s] + p[-(int)s - 1] +
               p[-(int)s - (int)r] + p[-(int)s - (int)r - 1];
    }
}

// General texel read in the reference's 3-D pool coordinates (used when the
// covering is 0 and the footprint may leave the brick): x, y clamp, z wraps.
template <int BYTES>
__device__ inline float texel(const uint8_t *pool, const gc_geometry *g, int px, int py, int pz)
{
    const uint32_t bs = g->brick_size;
    const int PX = g->cache_size[0] * bs, PY = g->cache_size[1] * bs, PZ = g->cache_size[2] * bs;
    px = px < 0 ? 0 : (px >= PX ? PX - 1 : px);
    py = py < 0 ? 0 : (py >= PY ? PY - 1 : py);
    pz = pz < 0 ? PZ - 1 : (pz >= PZ ? pz - PZ : pz);
    const uint32_t slot = (uint32_t)px / bs + g->cache_size[0] * ((uint32_t)py / bs + g->cache_size[1] * ((uint32_t)pz / bs));
    const uint64_t off = (uint64_t)slot * bs * bs * bs + (px % bs) + (uint64_t)bs * ((py % bs) + (uint64_t)bs * (pz % bs));
    if constexpr (BYTES == 1) return (float)pool[off];
    else if constexpr (BYTES == 2) return (float)reinterpret_cast<const uint16_t *>(pool)[off];
    else return reinterpret_cast<const float *>(pool)[off];
}

template <typename ACC>
struct Acc4 {
    ACC x, y, z, w;
};

__device__ inline uint32_t pack(float r, float g, float b, float a)
{
    // rgbaFloatToInt (RayCasters.cu:528-536)
    const uint32_t R = (uint32_t)(__saturatef(r) * 255.f), G = (uint32_t)(__saturatef(g) * 255.f),
                   B = (uint32_t)(__saturatef(b) * 255.f), A = (uint32_t)(__saturatef(a) * 255.f);
    return (A << 24) | (B << 16) | (G << 8) | R;
}
__device__ inline uint32_t pack(double r, double g, double b, double a)
{
    auto sat = [](double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); };
    const uint32_t R = (uint32_t)(sat(r) * 255.0), G = (uint32_t)(sat(g) * 255.0),
                   B = (uint32_t)(sat(b) * 255.0), A = (uint32_t)(sat(a) * 255.0);
    return (A << 24) | (B << 16) | (G << 8) | R;
}

__device__ inline uint32_t remap_tile(uint32_t bid, uint32_t nblocks)
{
    // Workgroup bid runs on XCD bid % 8; give each XCD a contiguous range.
    const uint32_t xcd = bid & 7u, k = bid >> 3;
    uint32_t off = 0;
    for (uint32_t y = 0; y < xcd; ++y) off += (nblocks + 7u - y) >> 3;
    return off + k;
}

__device__ inline uint64_t wave_sum(uint64_t v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int BYTES, typename ACC, int TW, int TH>
__global__ __launch_bounds__(TW *TH) void k_dvr(const DvrParams p)
{
    __shared__ LevelLds lv[GC_MAX_LEVELS];
    __shared__ float4 tfl[TF_LDS];
    const int tid = threadIdx.x;
    const gc_geometry *g = p.geom;
    if (tid < (int)p.nb_lods) {
        LevelLds &L = lv[tid];
        for (int a = 0; a < 3; ++a) {
            L.ior[a] = g->ior[tid][a];
            L.rdv[a] = g->rdv[tid][a];
            L.ldf[a] = (float)g->level_dim[tid][a];
            L.vvf[a] = (float)g->virt_vox[tid][a];
            L.inv[a] = p.inv_levels[tid][a];
        }
        L.lcoord = g->level_coord[tid];
        L.step = p.lod_step[tid];
        L.kexp = p.kexp[tid];
    }
    for (uint32_t i = tid; i < p.tf_size && i < (uint32_t)TF_LDS; i += TW * TH) tfl[i] = p.tf[i];
    __syncthreads();
    const float4 *TF = p.tf_size <= (uint32_t)TF_LDS ? tfl : p.tf;

    const uint32_t ntiles = p.tiles_x * p.tiles_y;
    const uint32_t tile = remap_tile(blockIdx.x, ntiles);
    const uint32_t x = (tile % p.tiles_x) * TW + (tid % TW);
    const uint32_t y = (tile / p.tiles_x) * TH + (tid / TW);
    const bool active = x < p.width && y < p.height;

    uint64_t n_samples = 0, n_visits = 0, n_unmapped = 0;
    const uint32_t W = p.width, H = p.height;
    const uint32_t bs = g->brick_size, cov = g->covering, inner = g->inner;
    const uint32_t cx_ = g->cache_size[0], cy_ = g->cache_size[1];
    const uint32_t MX = g->mrpd_dim[0], MY = g->mrpd_dim[1];
    const uint64_t bvox = (uint64_t)bs * bs * bs;
    const uint8_t *pool = static_cast<const uint8_t *>(p.pool);
    const float *M = p.m;
    const uint32_t nl = p.nb_lods;

    if (active) {
        // RayCasters.cu:57-84
        uint32_t xn = x + 1;
        if (xn > W) xn = x - 1;
        uint32_t yn = y + 1;
        if (yn > H) yn = y - 1;
        const float u = (((float)x + 0.5f) / (float)W) * 2.0f - 1.0f;
        const float un = (((float)xn + 0.5f) / (float)W) * 2.0f - 1.0f;
        const float v = (((float)y + 0.5f) / (float)H) * 2.0f - 1.0f;
        const float vn = (((float)yn + 0.5f) / (float)H) * 2.0f - 1.0f;
        const F3 origin = f3(M[3], M[7], M[11]);
        const F3 dir = mulm(M, normalize(f3(u, v, -2.f)));
        const F3 dxn = mulm(M, normalize(f3(un, v, -2.f)));
        const F3 dyn = mulm(M, normalize(f3(u, vn, -2.f)));
        const F3 dleft = normalize(cross(dir, dyn));
        const F3 dup = normalize(cross(dir, dxn));

        // intersectBox (RayCasters.cu:506-525)
        const F3 invr = f3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
        const F3 tbot = invr * (f3(p.bmin[0], p.bmin[1], p.bmin[2]) - origin);
        const F3 ttop = invr * (f3(p.bmax[0], p.bmax[1], p.bmax[2]) - origin);
        const F3 tmin = f3(fminf(ttop.x, tbot.x), fminf(ttop.y, tbot.y), fminf(ttop.z, tbot.z));
        const F3 tmax = f3(fmaxf(ttop.x, tbot.x), fmaxf(ttop.y, tbot.y), fmaxf(ttop.z, tbot.z));
        float tnear = fmaxf(fmaxf(tmin.x, tmin.y), fmaxf(tmin.x, tmin.z));
        const float tfar = fminf(fminf(tmax.x, tmax.y), fminf(tmax.x, tmax.z));
        const bool hit = tfar > tnear;

        const uint64_t pi = (uint64_t)y * W + x;
        tnear = fmaxf(tnear, p.tbuf[pi]);
        Acc4<ACC> fc;
        if constexpr (sizeof(ACC) == 8) {
            const double4 c = reinterpret_cast<const double4 *>(p.cstate)[pi];
            fc = Acc4<ACC>{c.x, c.y, c.z, c.w};
        } else {
            const float4 c = reinterpret_cast<const float4 *>(p.cstate)[pi];
            fc = Acc4<ACC>{c.x, c.y, c.z, c.w};
        }
        bool unmapped = false;
        if (!hit) {
            fc.x = fc.y = fc.z = (ACC)0;
        } else {
            if (tnear < 0.f) tnear = 0.f;
            float t = tnear;
            F3 pos = origin + dir * tnear;
            const float epx = fabsf(dot(dir - dxn, dleft) * tfar);
            const float epy = fabsf(dot(dir - dyn, dup) * tfar);
            // LOD state (compute_LOD, RayCasters.cu:551-567).  pixelDist grows
            // monotonically with t, so the chosen level never decreases along
            // the ray and the scan can resume from the current level.
            uint32_t lod = 0;
            float nsl = 0.f, nsu = 0.f;
            if (nl > 1) { nsl = spacing(dleft, lv[1].inv); nsu = spacing(dup, lv[1].inv); }
            uint32_t pcx = 0xffffffffu, pcy = 0xffffffffu, pcz = 0xffffffffu;
            uint4 e = make_uint4(0, 0, 0, 0);
            const uint8_t *bbase = pool;
            const ACC astop = (ACC)p.alpha_stop;
            const uint32_t N = p.tf_size;
            for (;;) {
                if (!(fc.w < astop && t < tfar)) break;
                const float al = t / tfar;
                const float pdx = epx * al, pdy = epy * al;
                F3 tp = pos * 0.495f + f3(0.495f, 0.495f, 0.495f);
                tp.z = 1.0f - tp.z;
                const F3 np = f3(fmaxf(0.f, fminf(tp.x, 0.99f)), fmaxf(0.f, fminf(tp.y, 0.99f)),
                                 fmaxf(0.f, fminf(tp.z, 0.99f)));
                while (lod + 1 < nl && !(nsl >= pdx * 1.0f || nsu >= pdy * 1.0f)) {
                    ++lod;
                    if (lod + 1 < nl) { nsl = spacing(dleft, lv[lod + 1].inv); nsu = spacing(dup, lv[lod + 1].inv); }
                }
                const LevelLds &L = lv[lod];
                const F3 pr = np * f3(L.ior[0], L.ior[1], L.ior[2]);
                const F3 pv = pr * f3(L.rdv[0], L.rdv[1], L.rdv[2]);
                const uint32_t bx = (uint32_t)(pv.x * L.ldf[0]);
                const uint32_t by = (uint32_t)(pv.y * L.ldf[1]);
                const uint32_t bz = (uint32_t)(pv.z * L.ldf[2]);
                const uint32_t mcx = L.lcoord + bx;
                if (mcx != pcx || by != pcy || bz != pcz) {
                    pcx = mcx; pcy = by; pcz = bz;
                    e = p.mrpd[mcx + (uint64_t)MX * (by + (uint64_t)MY * bz)];
                    ++n_visits;
                    if (e.w == 1) {
                        const uint32_t slot = e.x / bs + cx_ * (e.y / bs + cy_ * (e.z / bs));
                        p.usage[slot] = p.ts;
                        bbase = pool + (uint64_t)slot * bvox * BYTES;
                    }
                }
                if (e.w == 0) {
                    // raise_brick_request with the real-volume position
                    const uint32_t rx = L.lcoord + (uint32_t)(pr.x * L.ldf[0]);
                    const uint32_t ry = (uint32_t)(pr.y * L.ldf[1]);
                    const uint32_t rz = (uint32_t)(pr.z * L.ldf[2]);
                    p.request[rx + (uint64_t)MX * (ry + (uint64_t)MY * rz)] = p.ts;
                    p.tbuf[pi] = t;
                    if constexpr (sizeof(ACC) == 8)
                        reinterpret_cast<double4 *>(p.cstate)[pi] = make_double4(fc.x, fc.y, fc.z, fc.w);
                    else
                        reinterpret_cast<float4 *>(p.cstate)[pi] = make_float4(fc.x, fc.y, fc.z, fc.w);
                    unmapped = true;
                    ++n_unmapped;
                    break;
                }
                if (e.w != 1) {
                    // Empty brick: advance by tstep (RayCasters.cu:154-159)
                    t += p.tstep;
                    if (t > tfar) break;
                    pos = pos + dir * p.tstep;
                    continue;
                }
                // voxel within the brick: (v mod (bs-2c)) using the brick index
                // already computed (exact; falls back when FP32 rounding puts
                // the voxel index in a neighbouring brick).
                const uint32_t vx = (uint32_t)(pv.x * L.vvf[0]);
                const uint32_t vy = (uint32_t)(pv.y * L.vvf[1]);
                const uint32_t vz = (uint32_t)(pv.z * L.vvf[2]);
                uint32_t lx = vx - bx * inner, ly = vy - by * inner, lz = vz - bz * inner;
                if (lx >= inner) lx = vx % inner;
                if (ly >= inner) ly = vy % inner;
                if (lz >= inner) lz = vz % inner;
                ++n_samples;
                ACC s;
                if (BYTES != 4 && cov >= 1) {
                    const uint8_t *q = bbase + ((uint64_t)(cov + lx) + (uint64_t)bs * ((cov + ly) + (uint64_t)bs * (cov + lz))) * BYTES;
                    const uint32_t sum = box_sum_inner<BYTES == 4 ? 1 : BYTES>(q, bs);
                    if constexpr (sizeof(ACC) == 8) s = (double)sum / (BYTES == 1 ? 2040.0 : 524280.0);
                    else s = (float)sum / (BYTES == 1 ? 2040.0f : 524280.0f);
                } else {
                    // general footprint in pool coordinates
                    const int ex = (int)(cov + e.x + lx), ey = (int)(cov + e.y + ly), ez = (int)(cov + e.z + lz);
                    float fsum = 0.f;
                    double dsum = 0.0;
                    for (int k = 0; k < 2; ++k)
                        for (int j = 0; j < 2; ++j)
                            for (int i = 0; i < 2; ++i) {
                                const float tv = texel<BYTES>(pool, g, ex - 1 + i, ey - 1 + j, ez - 1 + k);
                                fsum += tv;
                                dsum += tv;
                            }
                    if constexpr (BYTES == 4) {
                        if constexpr (sizeof(ACC) == 8) s = dsum / 8.0;
                        else s = fsum / 8.0f;
                    } else {
                        if constexpr (sizeof(ACC) == 8) s = dsum / (BYTES == 1 ? 2040.0 : 524280.0);
                        else s = (float)dsum / (BYTES == 1 ? 2040.0f : 524280.0f);
                    }
                }
                // tex1D<float4>, linear filter, normalised, clamp
                const ACC xf = s * (ACC)N - (ACC)0.5;
                const ACC fl = floor(xf);
                const ACC a = xf - fl;
                int i0 = (int)fl, i1 = i0 + 1;
                i0 = i0 < 0 ? 0 : (i0 > (int)N - 1 ? (int)N - 1 : i0);
                i1 = i1 < 0 ? 0 : (i1 > (int)N - 1 ? (int)N - 1 : i1);
                const float4 T0 = TF[i0], T1 = TF[i1];
                const ACC b = (ACC)1 - a;
                Acc4<ACC> col{b * (ACC)T0.x + a * (ACC)T1.x, b * (ACC)T0.y + a * (ACC)T1.y,
                              b * (ACC)T0.z + a * (ACC)T1.z, b * (ACC)T0.w + a * (ACC)T1.w};
                // opacity correction 1 - (1-a)^(LODStep/tstep) (RayCasters.cu:165)
                if constexpr (sizeof(ACC) == 8) {
                    col.w = 1.0 - pow(1.0 - col.w, (double)L.kexp);
                } else {
                    const float base = 1.f - col.w;
                    col.w = 1.f - __builtin_amdgcn_exp2f(L.kexp * __builtin_amdgcn_logf(base));
                }
                if (col.w > (ACC)0.001f) {
                    col.x *= col.w; col.y *= col.w; col.z *= col.w;
                    const ACC wgt = (ACC)1 - fc.w;
                    const ACC tx = col.x * wgt, ty = col.y * wgt, tz = col.z * wgt, tw = col.w * wgt;
                    fc.x += tx; fc.y += ty; fc.z += tz; fc.w += tw;
                }
                t += L.step;
                pos = pos + dir * L.step;
            }
            if (!unmapped) {
                p.tbuf[pi] = t;
                if constexpr (sizeof(ACC) == 8)
                    reinterpret_cast<double4 *>(p.cstate)[pi] = make_double4(fc.x, fc.y, fc.z, fc.w);
                else
                    reinterpret_cast<float4 *>(p.cstate)[pi] = make_float4(fc.x, fc.y, fc.z, fc.w);
            }
        }
        p.pixels[pi] = pack(fc.x, fc.y, fc.z, fc.w);
    }

    if (p.stats) {
        const uint64_t s0 = wave_sum(n_samples), s1 = wave_sum(n_visits),
                       s2 = wave_sum(active ? 1ull : 0ull), s3 = wave_sum(n_unmapped);
        if ((tid & 63) == 0) {
            if (s0) atomicAdd((unsigned long long *)&p.stats->mapped_samples, (unsigned long long)s0);
            if (s1) atomicAdd((unsigned long long *)&p.stats->brick_visits, (unsigned long long)s1);
            if (s2) atomicAdd((unsigned long long *)&p.stats->pixels, (unsigned long long)s2);
            if (s3) atomicAdd((unsigned long long *)&p.stats->unmapped_rays, (unsigned long long)s3);
        }
    }
}

template <int BYTES, typename ACC>
hipError_t launch_tile(DvrParams p, uint32_t tile, hipStream_t stream)
{
    int tw, th;
    switch (tile) {
    case 0: tw = 8; th = 8; break;
    case 1: tw = 16; th = 8; break;
    case 2: tw = 32; th = 4; break;
    case 3: tw = 64; th = 1; break;
    default: return hipErrorInvalidValue;
    }
    p.tiles_x = (p.width + tw - 1) / tw;
    p.tiles_y = (p.height + th - 1) / th;
    const uint32_t nb = p.tiles_x * p.tiles_y;
    if (nb == 0) return hipSuccess;
    switch (tile) {
    case 0: k_dvr<BYTES, ACC, 8, 8><<<nb, 64, 0, stream>>>(p); break;
    case 1: k_dvr<BYTES, ACC, 16, 8><<<nb, 128, 0, stream>>>(p); break;
    case 2: k_dvr<BYTES, ACC, 32, 4><<<nb, 128, 0, stream>>>(p); break;
    case 3: k_dvr<BYTES, ACC, 64, 1><<<nb, 64, 0, stream>>>(p); break;
    }
    return hipGetLastError();
}

} // namespace

hipError_t launch_dvr(const DvrParams &p, uint32_t bytes, uint32_t precision, uint32_t tile,
                      hipStream_t stream)
{
    const bool d = precision == 64;
    switch (bytes) {
    case 1: return d ? launch_tile<1, double>(p, tile, stream) : launch_tile<1, float>(p, tile, stream);
    case 2: return d ? launch_tile<2, double>(p, tile, stream) : launch_tile<2, float>(p, tile, stream);
    case 4: return d ? launch_tile<4, double>(p, tile, stream) : launch_tile<4, float>(p, tile, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace gc
