// kernels_dvr.hip -- the DVR ray caster for gfx950 (RayCastDVRImpl,
// RayCasters.cu:36-198, with K_CacheManager::get_normalized,
// K_CacheManager.hpp:267-361, and K_CacheDevice::get_normalized,
// K_CacheDevice.hpp:191-201,316-334, inlined).
//
// Layout and mapping
//  * One lane per pixel.  A wave owns a 64-ray tile (8x8, 16x4, 32x2 or 64x1
//    pixels) so its rays walk the same bricks; a workgroup holds 4 waves on 4
//    consecutive tiles and shares the per-level tables and the transfer
//    function in LDS.
//  * Workgroup ids are remapped so that consecutive tiles land on the same
//    XCD (workgroups are dealt round-robin over the 8 XCDs): each XCD walks a
//    contiguous band of the image and keeps its bricks in its own L2.
//  * The pool is brick-major (slot * bs^3), so the 2x2x2 footprint of a
//    sample is four row pairs inside one brick, each fetched with one 8-byte
//    load.
//
// Exactness of the sample trace.  Ray set-up, positions, LOD choice and every
// page-table / voxel index are FP32 in the reference's operation order (no
// contraction, IEEE divide and square root).  Three rewrites keep that trace
// bit-for-bit:
//  * LOD: the reference evaluates pixelDist = endPixelDist * (t / tfar) at
//    every sample.  That predicate is monotone in t, so the level can only
//    grow along a ray; the first float t at which the current level stops
//    qualifying is found once per level by bisection over float bit patterns
//    and each sample compares t against it.
//  * Voxel reuse: consecutive samples in the same voxel of the same brick and
//    level read the same eight texels; their classified colour is reused.
//  * The page entry and the usage stamp are refreshed once per brick visit.
// Compositing is FP32 (precision 32: opacity correction with the hardware
// exp2/log2, FMA blending) or FP64 (precision 64).
#include "gc_internal.hpp"

namespace gc {

namespace {

constexpr int TF_LDS = 256;
constexpr int WAVES = 4;
constexpr float FINF = __builtin_huge_valf();

struct LevelLds {
    float ior[3], rdv[3], ldf[3], vvf[3], inv[3];
    uint32_t lcoord;
    float step, kexp;
    uint32_t unit_rdv;   // rdv == (1,1,1): the virtual position is the real one
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
__device__ inline float clamp099(float v) { return fmaxf(0.f, fminf(v, 0.99f)); }

// pixelDist at distance t: endPixelDist * (t / tfar) (RayCasters.cu:124-126)
__device__ inline float pixel_dist(float ep, float t, float tfar) { return ep * (t / tfar); }

// Smallest float t in (lo, hi] with pixel_dist(ep, t) > s, given
// pixel_dist(ep, lo) <= s < pixel_dist(ep, hi).  Bisection over the bit
// patterns of non-negative floats (monotone in value), after a bracket guess.
__device__ float first_above(float ep, float s, float tfar, float lo, float hi)
{
    // lo >= 0 (a -0.0 lower bound is read as +0.0)
    uint32_t lb = __float_as_uint(lo) & 0x7fffffffu, hb = __float_as_uint(hi);
    const float est = (s / ep) * tfar;
    if (est > lo && est < hi) {
        const uint32_t eb = __float_as_uint(est);
        const uint32_t w = 64;
        const uint32_t a = eb - lb > w ? eb - w : lb, b = hb - eb > w ? eb + w : hb;
        const bool a_ok = a == lb || pixel_dist(ep, __uint_as_float(a), tfar) <= s;
        const bool b_ok = b == hb || pixel_dist(ep, __uint_as_float(b), tfar) > s;
        if (a_ok && b_ok) { lb = a; hb = b; }
    }
    while (hb - lb > 1) {
        const uint32_t mb = lb + ((hb - lb) >> 1);
        if (pixel_dist(ep, __uint_as_float(mb), tfar) > s) hb = mb;
        else lb = mb;
    }
    return __uint_as_float(hb);
}

// First t >= t0 at which level `lod` stops being chosen: both pixel
// distances exceed the next level's spacings (compute_LOD,
// RayCasters.cu:551-567).  +inf if that never happens before tfar.
__device__ float level_switch(float t0, float tfar, float epx, float epy, float sl, float su)
{
    const bool x_far = pixel_dist(epx, tfar, tfar) > sl;
    const bool y_far = pixel_dist(epy, tfar, tfar) > su;
    if (!x_far || !y_far) return FINF;
    const float tx = pixel_dist(epx, t0, tfar) > sl ? t0 : first_above(epx, sl, tfar, t0, tfar);
    const float ty = pixel_dist(epy, t0, tfar) > su ? t0 : first_above(epy, su, tfar, t0, tfar);
    return fmaxf(tx, ty);
}

// Sum of the 8 texels {e-1, e}^3 of a voxel strictly inside its brick
// (covering >= 1).  Each (x-1, x) pair is taken from one 8-byte load at the
// 4-byte-aligned address below it (the pool is padded for the overrun).
template <int BYTES>
__device__ inline uint32_t box_sum_pairs(const uint8_t *brick, uint32_t bs, uint32_t ex, uint32_t ey, uint32_t ez)
{
    const uint64_t row = (uint64_t)bs * BYTES;
    const uint8_t *r0 = brick + ((uint64_t)(ez - 1) * bs + (ey - 1)) * row + (uint64_t)(ex - 1) * BYTES;
    const uint8_t *rows[4] = {r0, r0 + row, r0 + row * bs, r0 + row * bs + row};
    uint2 v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uintptr_t a = reinterpret_cast<uintptr_t>(rows[i]);
        __builtin_memcpy(&v[i], __builtin_assume_aligned(reinterpret_cast<const void *>(a & ~(uintptr_t)3), 4), 8);
    }
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t sh = (uint32_t)(reinterpret_cast<uintptr_t>(rows[i]) & 3) * 8;
        const uint64_t q = ((uint64_t)v[i].y << 32) | v[i].x;
        const uint32_t w = (uint32_t)(q >> sh);
        if constexpr (BYTES == 1) sum += (w & 0xffu) + ((w >> 8) & 0xffu);
        else sum += (w & 0xffffu) + (w >> 16);
    }
    return sum;
}

// General texel read in the reference's 3-D pool coordinates (covering 0, or
// float voxels): x, y clamp, z wraps (Texture3dDevice.hpp:104-122).
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

__device__ inline uint32_t remap_block(uint32_t bid, uint32_t nblocks)
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

// Classified, opacity-corrected, premultiplied colour of a sample value
// (tex1D linear + clamp, RayCasters.cu:163-175).  skip: alpha <= 0.001.
template <typename ACC>
__device__ inline Acc4<ACC> classify(ACC s, const float4 *TF, uint32_t N, float kexp, bool &skip)
{
    const ACC xf = s * (ACC)N - (ACC)0.5;
    const ACC fl = floor(xf);
    const ACC a = xf - fl;
    int i0 = (int)fl, i1 = i0 + 1;
    i0 = i0 < 0 ? 0 : (i0 > (int)N - 1 ? (int)N - 1 : i0);
    i1 = i1 < 0 ? 0 : (i1 > (int)N - 1 ? (int)N - 1 : i1);
    const float4 T0 = TF[i0], T1 = TF[i1];
    Acc4<ACC> col;
    if constexpr (sizeof(ACC) == 8) {
        const double b = 1.0 - a;
        col = Acc4<ACC>{b * (double)T0.x + a * (double)T1.x, b * (double)T0.y + a * (double)T1.y,
                        b * (double)T0.z + a * (double)T1.z, b * (double)T0.w + a * (double)T1.w};
        col.w = 1.0 - pow(1.0 - col.w, (double)kexp);
    } else {
        col = Acc4<ACC>{fmaf(a, T1.x - T0.x, T0.x), fmaf(a, T1.y - T0.y, T0.y), fmaf(a, T1.z - T0.z, T0.z),
                        fmaf(a, T1.w - T0.w, T0.w)};
        col.w = 1.f - __builtin_amdgcn_exp2f(kexp * __builtin_amdgcn_logf(1.f - col.w));
    }
    skip = !(col.w > (ACC)0.001f);
    col.x *= col.w;
    col.y *= col.w;
    col.z *= col.w;
    return col;
}

template <int BYTES, typename ACC, int TW, int TH>
__global__ __launch_bounds__(64 * WAVES) void k_dvr(const DvrParams p)
{
    static_assert(TW * TH == 64, "one wave per ray tile");
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
        L.unit_rdv = L.rdv[0] == 1.f && L.rdv[1] == 1.f && L.rdv[2] == 1.f;
    }
    for (uint32_t i = tid; i < p.tf_size && i < (uint32_t)TF_LDS; i += 64 * WAVES) tfl[i] = p.tf[i];
    __syncthreads();
    const float4 *TF = p.tf_size <= (uint32_t)TF_LDS ? tfl : p.tf;

    const uint32_t ntiles = p.tiles_x * p.tiles_y;
    const uint32_t nblocks = (ntiles + WAVES - 1) / WAVES;
    const uint32_t tile = remap_block(blockIdx.x, nblocks) * WAVES + (tid >> 6);
    const uint32_t lane = tid & 63;
    const uint32_t x = (tile % p.tiles_x) * TW + (lane % TW);
    const uint32_t y = (tile / p.tiles_x) * TH + (lane / TW);
    const bool active = tile < ntiles && x < p.width && y < p.height;

    uint64_t n_samples = 0, n_visits = 0, n_unmapped = 0;
    const uint32_t W = p.width, H = p.height;
    const uint32_t bs = g->brick_size, cov = g->covering, inner = g->inner;
    const uint32_t cx_ = g->cache_size[0], cy_ = g->cache_size[1];
    const uint32_t MX = g->mrpd_dim[0], MY = g->mrpd_dim[1];
    const uint64_t bbytes = (uint64_t)bs * bs * bs * BYTES;
    const uint8_t *pool = static_cast<const uint8_t *>(p.pool);
    const float *M = p.m;
    const uint32_t nl = p.nb_lods;
    const bool fast_fetch = BYTES != 4 && cov >= 1;
    const ACC inv_norm = BYTES == 1 ? (ACC)(1.0 / 2040.0) : (ACC)(1.0 / 524280.0);

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
            const ACC astop = (ACC)p.alpha_stop;
            const uint32_t N = p.tf_size;

            // level state: lod, the distance where it stops qualifying, and
            // the level's constants
            uint32_t lod = 0;
            float tswitch = -1.f;   // forces a settle before the first sample
            float step = 0.f, kexp = 0.f;
            F3 dstep = f3(0.f, 0.f, 0.f);
            // brick / voxel caches
            uint32_t pcx = 0xffffffffu, pcy = 0xffffffffu, pcz = 0xffffffffu;
            uint4 e = make_uint4(0, 0, 0, 0);
            const uint8_t *bbase = pool;
            uint32_t pvx = 0xffffffffu, pvy = 0, pvz = 0;
            Acc4<ACC> cc{0, 0, 0, 0};
            bool cskip = true;

            for (;;) {
                if (!(fc.w < astop && t < tfar)) break;
                if (t >= tswitch) {
                    // settle the level at t exactly as compute_LOD does, then
                    // find where it next changes
                    const float pdx = pixel_dist(epx, t, tfar), pdy = pixel_dist(epy, t, tfar);
                    while (lod + 1 < nl &&
                           !(spacing(dleft, lv[lod + 1].inv) >= pdx * 1.0f || spacing(dup, lv[lod + 1].inv) >= pdy * 1.0f))
                        ++lod;
                    tswitch = lod + 1 < nl ? level_switch(t, tfar, epx, epy, spacing(dleft, lv[lod + 1].inv),
                                                          spacing(dup, lv[lod + 1].inv))
                                           : FINF;
                    step = lv[lod].step;
                    kexp = lv[lod].kexp;
                    dstep = dir * step;
                }
                const LevelLds &L = lv[lod];
                F3 tp = pos * 0.495f + f3(0.495f, 0.495f, 0.495f);
                tp.z = 1.0f - tp.z;
                const F3 np = f3(clamp099(tp.x), clamp099(tp.y), clamp099(tp.z));
                const F3 pr = np * f3(L.ior[0], L.ior[1], L.ior[2]);
                const F3 pv = L.unit_rdv ? pr : pr * f3(L.rdv[0], L.rdv[1], L.rdv[2]);
                const uint32_t bx = (uint32_t)(pv.x * L.ldf[0]);
                const uint32_t by = (uint32_t)(pv.y * L.ldf[1]);
                const uint32_t bz = (uint32_t)(pv.z * L.ldf[2]);
                const uint32_t mcx = L.lcoord + bx;
                if (mcx != pcx || by != pcy || bz != pcz) {
                    pcx = mcx; pcy = by; pcz = bz;
                    e = p.mrpd[mcx + (uint64_t)MX * (by + (uint64_t)MY * bz)];
                    ++n_visits;
                    pvx = 0xffffffffu;
                    if (e.w == 1) {
                        const uint32_t slot = e.x / bs + cx_ * (e.y / bs + cy_ * (e.z / bs));
                        p.usage[slot] = p.ts;
                        bbase = pool + (uint64_t)slot * bbytes;
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
                const uint32_t vx = (uint32_t)(pv.x * L.vvf[0]);
                const uint32_t vy = (uint32_t)(pv.y * L.vvf[1]);
                const uint32_t vz = (uint32_t)(pv.z * L.vvf[2]);
                if (vx != pvx || vy != pvy || vz != pvz) {
                    pvx = vx; pvy = vy; pvz = vz;
                    // voxel within the brick: v mod (bs-2c), via the brick
                    // index already known (falls back when FP32 rounding put
                    // the voxel index in a neighbouring brick)
                    uint32_t lx = vx - bx * inner, ly = vy - by * inner, lz = vz - bz * inner;
                    if (lx >= inner) lx = vx % inner;
                    if (ly >= inner) ly = vy % inner;
                    if (lz >= inner) lz = vz % inner;
                    ACC s;
                    if (fast_fetch) {
                        const uint32_t sum = box_sum_pairs<BYTES == 4 ? 1 : BYTES>(bbase, bs, cov + lx, cov + ly, cov + lz);
                        if constexpr (sizeof(ACC) == 8) s = (double)sum / (BYTES == 1 ? 2040.0 : 524280.0);
                        else s = (float)sum * inv_norm;
                    } else {
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
                            else s = (float)dsum * inv_norm;
                        }
                    }
                    cc = classify<ACC>(s, TF, N, kexp, cskip);
                }
                ++n_samples;
                if (!cskip) {
                    // front-to-back blending (RayCasters.cu:176-178)
                    const ACC wgt = (ACC)1 - fc.w;
                    if constexpr (sizeof(ACC) == 8) {
                        const ACC tx = cc.x * wgt, ty = cc.y * wgt, tz = cc.z * wgt, tw = cc.w * wgt;
                        fc.x += tx; fc.y += ty; fc.z += tz; fc.w += tw;
                    } else {
                        fc.x = fmaf(cc.x, wgt, fc.x);
                        fc.y = fmaf(cc.y, wgt, fc.y);
                        fc.z = fmaf(cc.z, wgt, fc.z);
                        fc.w = fmaf(cc.w, wgt, fc.w);
                    }
                }
                t += step;
                pos = pos + dstep;
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
    static const int TW[4] = {8, 16, 32, 64};
    if (tile > 3) return hipErrorInvalidValue;
    const int tw = TW[tile], th = 64 / tw;
    p.tiles_x = (p.width + tw - 1) / tw;
    p.tiles_y = (p.height + th - 1) / th;
    const uint32_t ntiles = p.tiles_x * p.tiles_y;
    const uint32_t nb = (ntiles + WAVES - 1) / WAVES;
    if (nb == 0) return hipSuccess;
    switch (tile) {
    case 0: k_dvr<BYTES, ACC, 8, 8><<<nb, 64 * WAVES, 0, stream>>>(p); break;
    case 1: k_dvr<BYTES, ACC, 16, 4><<<nb, 64 * WAVES, 0, stream>>>(p); break;
    case 2: k_dvr<BYTES, ACC, 32, 2><<<nb, 64 * WAVES, 0, stream>>>(p); break;
    case 3: k_dvr<BYTES, ACC, 64, 1><<<nb, 64 * WAVES, 0, stream>>>(p); break;
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
