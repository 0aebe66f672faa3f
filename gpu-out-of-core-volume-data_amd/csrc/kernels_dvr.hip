// kernels_dvr.hip -- the DVR ray caster for gfx950 (RayCastDVRImpl,
// RayCasters.cu:36-198, with K_CacheManager::get_normalized,
// K_CacheManager.hpp:267-361, and K_CacheDevice::get_normalized,
// K_CacheDevice.hpp:191-201,316-334, inlined).
//
// Work distribution
//  * Rays come in packets of 64: 64 entries of an explicit ray list (the
//    rays a previous launch stopped at an Unmapped brick, view << 26 | pixel)
//    come first, then the TWxTH pixel tiles (8x8, 16x4, 32x2 or 64x1) of the
//    views that start in this launch.  One launch can cover many views
//    (batched rendering) and mix continuing and starting views (serving).
//  * The kernel is persistent: waves pull packets from 8 per-XCD queues
//    (workgroup b runs on XCD b % 8 and starts on that XCD's contiguous share
//    of the packets, so neighbouring tiles share an L2), stealing from the
//    other queues when their own is empty.
//  * A lane whose ray ends is refilled with the next ray of the wave's packet
//    stream once enough lanes are idle, so long and short rays do not hold a
//    wave hostage.
//
// Per-sample cost.  The march loop is issue-bound, so everything that does
// not change from one sample to the next is hoisted out of it:
//  * the parameters of the ray's current level live in registers and are
//    reloaded from LDS only when the level changes;
//  * the distances at which a ray changes level are computed once, at ray
//    set-up, into a per-lane LDS table (no search inside the march);
//  * the page entry is read with one 4-byte load per brick visit from the
//    cell table (slot << 2 | state per page-directory cell, kept equal to
//    the 16-byte MRPD entries by the kernels that write them), which also
//    yields the brick's base address in the brick-major pool;
//  * the 2x2x2 footprint of a sample is four texel pairs (x-1, x) in one
//    brick: four 4-byte (u16) or 2-byte (u8) loads at 2- or 1-byte alignment,
//    summed with v_sad_u16 / v_sad_u8;
//  * consecutive samples of a ray in the same voxel (steps are half a voxel of
//    the current level) reuse the box sum: those lanes issue no loads, which
//    is what bounds close-up views, where the memory pipeline saturates.
//
// Software pipelining (u8/u16 voxels with covering >= 1, realDivVirtual == 1,
// the configurations every bricked level gives).  Each ray keeps three
// samples in flight: the page entry of sample j+2 and the footprint of sample
// j+1 are loading while sample j is classified and composited.  Every
// half-step issues exactly one entry load and four footprint loads, for every
// lane (lanes that need none load one wave-uniform address), so the
// compiler's in-order waits land one half-step after each load instead of
// right after it.  Side effects (usage/request stamps, counters, stored
// state) happen only when a sample is consumed, in the reference's order.
//
// Exactness of the sample trace.  Ray set-up, positions, LOD choice and every
// page-table / voxel index are FP32 in the reference's operation order (no
// contraction, IEEE divide and square root).  The LOD rewrite keeps that trace
// bit-for-bit: the reference evaluates pixelDist = endPixelDist * (t / tfar)
// at every sample and picks the coarsest level whose voxel spacing both pixel
// distances exceed.  pixelDist is monotone in t and the spacings are monotone
// in the level, so level(t) = max{l : T_l <= t} where T_l is the first float
// t at which level l qualifies; T_l is found once per level by bisection over
// float bit patterns.  Rays are independent, so processing order (packets,
// refills, ray lists) does not change any pixel, usage stamp or request.
// Compositing is FP32 (precision 32: opacity correction with the hardware
// exp2/log2, FMA blending) or FP64 (precision 64).
#include "gc_internal.hpp"

#include <algorithm>

namespace gc {

namespace {

constexpr int WAVES = 4;
constexpr int LANES = 64 * WAVES;
constexpr float FINF = __builtin_huge_valf();
constexpr uint32_t NONE = 0xffffffffu;
// Waves per SIMD the FP32 tile march is compiled for (its register budget).
#ifndef DVR_WAVES_PER_EU
#define DVR_WAVES_PER_EU 4
#endif

struct LevelLds {
    float ior[3], rdv[3], ldf[3], vvf[3], inv[3];
    uint32_t lcoord;
    float step, kexp;
    int32_t esq;        // exact opacity factor: squarings (see opacity_factor)
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
    const float inv = 1.0f / sqrtf(dot(v, v));   // correctly rounded (build flag)
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
    return sqrtf(dot(a, a));
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
        // brackets of +-4 ulp, then +-64 ulp, around the closed-form estimate;
        // each is used only when the predicate confirms it, so the result is
        // the bisection's over the full range
        const uint32_t eb = __float_as_uint(est);
        for (uint32_t w = 4; w <= 64; w *= 16) {
            const uint32_t a = eb - lb > w ? eb - w : lb, b = hb - eb > w ? eb + w : hb;
            const bool a_ok = a == lb || pixel_dist(ep, __uint_as_float(a), tfar) <= s;
            const bool b_ok = b == hb || pixel_dist(ep, __uint_as_float(b), tfar) > s;
            if (a_ok && b_ok) {
                lb = a;
                hb = b;
                break;
            }
        }
    }
    while (hb - lb > 1) {
        const uint32_t mb = lb + ((hb - lb) >> 1);
        if (pixel_dist(ep, __uint_as_float(mb), tfar) > s) hb = mb;
        else lb = mb;
    }
    return __uint_as_float(hb);
}

// First t >= t0 at which a level with voxel spacings (sl, su) qualifies:
// both pixel distances exceed them (compute_LOD, RayCasters.cu:551-567).
// +inf if that never happens before tfar.
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
// (covering >= 1).  `a` addresses texel (ex-1, ey-1, ez-1); each (x-1, x)
// pair is one unaligned load.
template <int BYTES>
__device__ inline uint32_t box_sum(const uint8_t *a, uint32_t row, uint32_t plane)
{
    if constexpr (BYTES == 1) {
        uint16_t h0, h1, h2, h3;
        __builtin_memcpy(&h0, a, 2);
        __builtin_memcpy(&h1, a + row, 2);
        __builtin_memcpy(&h2, a + plane, 2);
        __builtin_memcpy(&h3, a + plane + row, 2);
        uint32_t s = __builtin_amdgcn_sad_u8(h0, 0u, 0u);
        s = __builtin_amdgcn_sad_u8(h1, 0u, s);
        s = __builtin_amdgcn_sad_u8(h2, 0u, s);
        return __builtin_amdgcn_sad_u8(h3, 0u, s);
    } else {
        uint32_t w0, w1, w2, w3;
        __builtin_memcpy(&w0, a, 4);
        __builtin_memcpy(&w1, a + row, 4);
        __builtin_memcpy(&w2, a + plane, 4);
        __builtin_memcpy(&w3, a + plane + row, 4);
        uint32_t s = __builtin_amdgcn_sad_u16(w0, 0u, 0u);
        s = __builtin_amdgcn_sad_u16(w1, 0u, s);
        s = __builtin_amdgcn_sad_u16(w2, 0u, s);
        return __builtin_amdgcn_sad_u16(w3, 0u, s);
    }
}

// The same footprint as four raw words, loaded without waiting for them (the
// pipelined march issues them one sample ahead of their use).
struct Quad {
    uint32_t a, b, c, d;
};
template <int BYTES>
__device__ inline Quad box_load(const uint8_t *a, uint32_t row, uint32_t plane)
{
    Quad q;
    if constexpr (BYTES == 1) {
        uint16_t h0, h1, h2, h3;
        __builtin_memcpy(&h0, a, 2);
        __builtin_memcpy(&h1, a + row, 2);
        __builtin_memcpy(&h2, a + plane, 2);
        __builtin_memcpy(&h3, a + plane + row, 2);
        q = Quad{h0, h1, h2, h3};
    } else {
        __builtin_memcpy(&q.a, a, 4);
        __builtin_memcpy(&q.b, a + row, 4);
        __builtin_memcpy(&q.c, a + plane, 4);
        __builtin_memcpy(&q.d, a + plane + row, 4);
    }
    return q;
}
template <int BYTES>
__device__ inline uint32_t box_add(const Quad &q)
{
    if constexpr (BYTES == 1) {
        uint32_t s = __builtin_amdgcn_sad_u8(q.a, 0u, 0u);
        s = __builtin_amdgcn_sad_u8(q.b, 0u, s);
        s = __builtin_amdgcn_sad_u8(q.c, 0u, s);
        return __builtin_amdgcn_sad_u8(q.d, 0u, s);
    } else {
        uint32_t s = __builtin_amdgcn_sad_u16(q.a, 0u, 0u);
        s = __builtin_amdgcn_sad_u16(q.b, 0u, s);
        s = __builtin_amdgcn_sad_u16(q.c, 0u, s);
        return __builtin_amdgcn_sad_u16(q.d, 0u, s);
    }
}

// Views keep their buffer pointers in LDS; accesses through them are made
// explicitly global (a flat access would make every later wait on vector
// memory wait for everything in flight).
typedef float gf4 __attribute__((ext_vector_type(4)));
typedef double gd4 __attribute__((ext_vector_type(4)));
template <typename T>
__device__ inline __attribute__((address_space(1))) T *gp(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ inline const __attribute__((address_space(1))) T *gp(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
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

__device__ inline uint32_t wave_sum(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Opacity correction and premultiplication of an interpolated TF sample
// (RayCasters.cu:163-175).  skip: alpha <= 0.001.
template <typename ACC, typename T4>
__device__ inline Acc4<ACC> shade(ACC a, T4 T0, T4 T1, float kexp, bool &skip)
{
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

// tex1D with linear filtering and clamp addressing over N entries.
template <typename ACC, typename TFP>
__device__ inline Acc4<ACC> classify(ACC s, TFP TF, uint32_t N, float kexp, bool &skip)
{
    const ACC xf = s * (ACC)N - (ACC)0.5;
    const ACC fl = floor(xf);
    const ACC a = xf - fl;
    int i0 = (int)fl, i1 = i0 + 1;
    i0 = i0 < 0 ? 0 : (i0 > (int)N - 1 ? (int)N - 1 : i0);
    i1 = i1 < 0 ? 0 : (i1 > (int)N - 1 ? (int)N - 1 : i1);
    return shade<ACC>(a, TF[i0], TF[i1], kexp, skip);
}

// The same lookup in the LDS copy of the table, which carries one clamped
// entry on each side (TFp[j] = TF[clamp(j - 1)]).  Valid for s in [0, 1]
// (integer voxels), where floor(s * N - 0.5) lies in [-1, N - 1].
template <typename ACC>
__device__ inline Acc4<ACC> classify_padded(ACC s, const float4 *TFp, uint32_t N, float kexp, bool &skip)
{
    const ACC xf = s * (ACC)N - (ACC)0.5;
    const ACC fl = floor(xf);
    const ACC a = xf - fl;
    const int j = (int)fl + 1;
    return shade<ACC>(a, TFp[j], TFp[j + 1], kexp, skip);
}

// Texture coordinate of an integer box sum.  The sample is s = sum / D
// (D = 8 * 255 for u8, 8 * 65535 for u16) and the coordinate xf = s * N - 0.5
// (RayCasters.cu:163); xf + 1 = (2 sum N + D) / (2 D), so q = floor(xf) + 1
// and the lerp weight frac(xf) = rem / (2 D) follow from integer arithmetic.
// In FP32, forming s and xf as floats would round xf to 2^-15 at N = 256, an
// error in the weight that the lerp carries into every colour and alpha.
// WIDE: 64-bit numerator (tables above 4095 entries); else 2 sum N + D < 2^32.
template <int BYTES, bool WIDE>
__device__ inline void tf_coord(uint32_t sum, uint32_t N, uint32_t &q, uint32_t &rem)
{
    constexpr uint32_t D = BYTES == 1 ? 2040u : 524280u;
    if constexpr (WIDE) {
        const uint64_t num = 2ull * sum * N + D;
        const uint64_t qq = num / (2u * D);
        q = (uint32_t)qq;
        rem = (uint32_t)(num - qq * (2u * D));
    } else {
        const uint32_t num = 2u * sum * N + D;
        q = num / (2u * D);
        rem = num - q * (2u * D);
    }
}

// FP32 classification of an integer box sum at the exact coordinate.  It
// also returns x = 1 - a in double (a: the interpolated transfer-function
// alpha, from the double weight), the input of the exact opacity decisions.
template <int BYTES, typename T4>
__device__ inline Acc4<float> shade_exact(uint32_t rem, T4 T0, T4 T1, float kexp, bool &skip, double &x)
{
    constexpr uint32_t D2 = 2u * (BYTES == 1 ? 2040u : 524280u);
    const float a = (float)rem * (1.0f / (float)D2);
    const double wd = (double)rem * (1.0 / (double)D2);
    x = 1.0 - fma(wd, (double)T1.w - (double)T0.w, (double)T0.w);
    return shade<float>(a, T0, T1, kexp, skip);
}

// Classification of an integer box sum.  FP64: as the FP64 caster computes
// it (s = sum / D, then s * N - 0.5), skip from the FP64 alpha.  FP32: the
// exact coordinate, skip from the FP32 alpha, and x as above.  Padded LDS
// table (N <= DVR_TF_LDS_MAX).
template <int BYTES, typename ACC>
__device__ inline Acc4<ACC> classify_sum_padded(uint32_t sum, const float4 *TFp, uint32_t N, float kexp, bool &skip,
                                                double &x)
{
    if constexpr (sizeof(ACC) == 8) {
        return classify_padded<double>((double)sum / (BYTES == 1 ? 2040.0 : 524280.0), TFp, N, kexp, skip);
    } else {
        uint32_t q, rem;
        tf_coord<BYTES, false>(sum, N, q, rem);
        return shade_exact<BYTES>(rem, TFp[q], TFp[q + 1], kexp, skip, x);
    }
}
// Table in HBM (any N): entries clamp(q - 1) and clamp(q).
template <int BYTES, typename ACC, typename TFP>
__device__ inline Acc4<ACC> classify_sum(uint32_t sum, TFP TF, uint32_t N, float kexp, bool &skip, double &x)
{
    if constexpr (sizeof(ACC) == 8) {
        return classify<double>((double)sum / (BYTES == 1 ? 2040.0 : 524280.0), TF, N, kexp, skip);
    } else {
        uint32_t q, rem;
        tf_coord<BYTES, true>(sum, N, q, rem);
        const uint32_t i0 = q == 0 ? 0u : (q - 1u > N - 1u ? N - 1u : q - 1u);
        const uint32_t i1 = q > N - 1u ? N - 1u : q;
        return shade_exact<BYTES>(rem, TF[i0], TF[i1], kexp, skip, x);
    }
}
// Float voxels (FP32 colour from the float sample, x from the double one).
template <typename TFP>
__device__ inline Acc4<float> classify_float_x(float s, double sd, TFP TF, uint32_t N, float kexp, bool &skip,
                                               double &x)
{
    const double xf = sd * (double)N - 0.5;
    const double fl = floor(xf);
    const double ad = xf - fl;
    int i0 = (int)fl, i1 = i0 + 1;
    i0 = i0 < 0 ? 0 : (i0 > (int)N - 1 ? (int)N - 1 : i0);
    i1 = i1 < 0 ? 0 : (i1 > (int)N - 1 ? (int)N - 1 : i1);
    x = 1.0 - ((1.0 - ad) * (double)TF[i0].w + ad * (double)TF[i1].w);
    return classify<float>(s, TF, N, kexp, skip);
}

// Exact opacity decisions of FP32 compositing.  The FP64 caster keeps alpha
// as fc.w += a'(1 - fc.w) with a' = 1 - x^k (k the level's opacity-correction
// exponent), skips samples with a' <= 0.001 and stops once fc.w >= the alpha
// stop (RayCasters.cu:122,165-178).  The transmittance 1 - fc.w is the
// product of x^k over the blended samples, i.e. Q^(k0) with Q the product of
// y = x^(k / k0) (k0: level 0's exponent).  Accumulating Q in double makes
// every skip and stop decision the FP64 caster's (their thresholds become
// (1 - 0.001)^(1/k0) and (1 - alpha stop)^(1/k0)), so FP32 compositing
// follows the FP64 sample trace; only the colours are FP32.  y is x squared
// esq times, k / k0 = 2^esq: every level of a volume whose smallest extent
// halves exactly from level to level (power-of-two volumes).  Other volumes
// (DvrLaunch::exact_alpha == 0) take the FP32 decisions.
__device__ inline double opacity_factor(double x, int esq)
{
    for (int i = 0; i < esq; ++i) x *= x;
    return x;
}

// Next packet for a wave on XCD `xcd`: its own contiguous share first, then
// the other XCDs' shares.  NONE when every queue is drained.
__device__ inline uint32_t fetch_packet(const DvrLaunch &p, uint32_t n_packets, uint32_t xcd)
{
    for (uint32_t k = 0; k < 8; ++k) {
        const uint32_t x = (xcd + k) & 7u;
        const uint32_t lo = (uint32_t)((uint64_t)n_packets * x / 8);
        const uint32_t hi = (uint32_t)((uint64_t)n_packets * (x + 1) / 8);
        if (lo >= hi) continue;
        const uint32_t c = atomicAdd(&p.queue[x], 1u);
        if (lo + c < hi) return lo + c;
    }
    return NONE;
}

// Ray set-up of pixel (x, y) of a view (RayCasters.cu:57-120): eye ray, the
// neighbouring rays that give the pixel footprint, and the box interval.
struct RaySetup {
    F3 origin, dir, dxn, dyn, dleft, dup;
    float tnear, tfar;
    bool hit;
};
__device__ inline RaySetup setup_ray(const DvrView &V, uint32_t x, uint32_t y)
{
    RaySetup r;
    const uint32_t W = V.width, H = V.height;
    uint32_t xn = x + 1;
    if (xn > W) xn = x - 1;
    uint32_t yn = y + 1;
    if (yn > H) yn = y - 1;
    const float u = (((float)x + 0.5f) / (float)W) * 2.0f - 1.0f;
    const float un = (((float)xn + 0.5f) / (float)W) * 2.0f - 1.0f;
    const float vy = (((float)y + 0.5f) / (float)H) * 2.0f - 1.0f;
    const float vn = (((float)yn + 0.5f) / (float)H) * 2.0f - 1.0f;
    const float *M = V.m;
    r.origin = f3(M[3], M[7], M[11]);
    r.dir = mulm(M, normalize(f3(u, vy, -2.f)));
    r.dxn = mulm(M, normalize(f3(un, vy, -2.f)));
    r.dyn = mulm(M, normalize(f3(u, vn, -2.f)));
    r.dleft = normalize(cross(r.dir, r.dyn));
    r.dup = normalize(cross(r.dir, r.dxn));
    // intersectBox (RayCasters.cu:506-525)
    const F3 invr = f3(1.0f / r.dir.x, 1.0f / r.dir.y, 1.0f / r.dir.z);
    const F3 tbot = invr * (f3(V.bmin[0], V.bmin[1], V.bmin[2]) - r.origin);
    const F3 ttop = invr * (f3(V.bmax[0], V.bmax[1], V.bmax[2]) - r.origin);
    const F3 tmin = f3(fminf(ttop.x, tbot.x), fminf(ttop.y, tbot.y), fminf(ttop.z, tbot.z));
    const F3 tmax = f3(fmaxf(ttop.x, tbot.x), fmaxf(ttop.y, tbot.y), fmaxf(ttop.z, tbot.z));
    r.tnear = fmaxf(fmaxf(tmin.x, tmin.y), fmaxf(tmin.x, tmin.z));
    r.tfar = fminf(fminf(tmax.x, tmax.y), fminf(tmax.x, tmax.z));
    r.hit = r.tfar > r.tnear;
    return r;
}

// Page cell and in-brick voxel of a sample position at one level
// (get_normalized, K_CacheManager.hpp:267-361, with realDivVirtual == 1; the
// MRPD x and y*z extents are below 2^24, checked on the host).
__device__ inline void locate_cell(F3 pos, F3 ior, F3 ldf, F3 vvf, uint32_t lcoord, uint32_t MX, uint32_t MY,
                                   uint32_t inner, uint32_t bs, uint32_t &cell, uint32_t &vi)
{
    F3 tp = pos * 0.495f + f3(0.495f, 0.495f, 0.495f);
    tp.z = 1.0f - tp.z;
    const F3 np = f3(clamp099(tp.x), clamp099(tp.y), clamp099(tp.z));
    const F3 pr = np * ior;
    const uint32_t bx = (uint32_t)(pr.x * ldf.x);
    const uint32_t by = (uint32_t)(pr.y * ldf.y);
    const uint32_t bz = (uint32_t)(pr.z * ldf.z);
    cell = lcoord + bx + __umul24(MX, by + __umul24(MY, bz));
    const uint32_t vx = (uint32_t)(pr.x * vvf.x);
    const uint32_t vy = (uint32_t)(pr.y * vvf.y);
    const uint32_t vz = (uint32_t)(pr.z * vvf.z);
    uint32_t lx = vx - __umul24(bx, inner), ly = vy - __umul24(by, inner), lz = vz - __umul24(bz, inner);
    if (__builtin_expect(lx >= inner || ly >= inner || lz >= inner, 0)) {
        lx = vx % inner;
        ly = vy % inner;
        lz = vz % inner;
    }
    vi = __umul24(__umul24(lz, bs) + ly, bs) + lx;
}

// Pool slot of a page entry (entry xyz = the slot's voxel origin in the
// reference's 3-D pool; coordinates are multiples of bs below 2^24).
// Slot and state of a cell-table word (cell_word: slot << 2 | w).
__device__ inline uint32_t word_slot(uint32_t c) { return c >> 2; }
__device__ inline uint32_t word_state(uint32_t c) { return c & 3u; }

// Per level, what the walk through the table caches needs (LDS copy).
struct PtLds {
    float tdf[3];                       // MRPD entries per axis
    uint32_t tcoord;                    // MRPD x offset of the level
    float tvf[GC_MAX_TABLES][3];        // per table: virtualized entries per axis
};

__device__ inline void pt_lds_fill(PtLds &Q, const gc_geometry *g, uint32_t l)
{
    for (int a = 0; a < 3; ++a) {
        Q.tdf[a] = (float)g->top_dim[l][a];
        for (int i = 0; i < GC_MAX_TABLES; ++i) Q.tvf[i][a] = (float)g->table[i].virt[l][a];
    }
    Q.tcoord = g->top_coord[l];
}

// The key of the entries a walk reads at the page position pv: the MRPD
// cell, then per table the entry inside its block, (uint)(pv * virtualized)
// % BlockSize (K_MultiResolutionPageDirectory.hpp:143-155,
// K_CacheDevice.hpp:316-334).  Positions with equal keys walk to the same
// entry.
struct PtKey {
    uint32_t top, loc[GC_MAX_TABLES];
};
// bx, by, bz: the last table's index before the modulo, i.e. the brick
// index (not part of the key)
__device__ inline PtKey pt_key(const PageTables &pt, const PtLds &Q, F3 pv, uint32_t &bx, uint32_t &by,
                               uint32_t &bz)
{
    PtKey k;
    k.top = Q.tcoord + (uint32_t)(pv.x * Q.tdf[0]) +
            pt.top_x * ((uint32_t)(pv.y * Q.tdf[1]) + pt.top_y * (uint32_t)(pv.z * Q.tdf[2]));
#pragma unroll
    for (int i = 0; i < GC_MAX_TABLES; ++i) {
        k.loc[i] = 0u;
        if ((uint32_t)i < pt.n) {
            const uint32_t ix = (uint32_t)(pv.x * Q.tvf[i][0]), iy = (uint32_t)(pv.y * Q.tvf[i][1]),
                           iz = (uint32_t)(pv.z * Q.tvf[i][2]);
            const uint32_t Bx = pt.block[i][0], By = pt.block[i][1], Bz = pt.block[i][2];
            uint32_t lx, ly, lz;
            if (pt.pow2[i]) {
                lx = ix & (Bx - 1u);
                ly = iy & (By - 1u);
                lz = iz & (Bz - 1u);
            } else {
                lx = ix % Bx;
                ly = iy % By;
                lz = iz % Bz;
            }
            k.loc[i] = lx + Bx * (ly + By * lz);
            if ((uint32_t)i + 1 == pt.n) {
                bx = ix;
                by = iy;
                bz = iz;
            }
        }
    }
    return k;
}
__device__ inline bool pt_same(const PtKey &a, const PtKey &b)
{
    bool e = a.top == b.top;
#pragma unroll
    for (int i = 0; i < GC_MAX_TABLES; ++i) e = e && a.loc[i] == b.loc[i];
    return e;
}
// The walk of get_normalized (K_CacheManager.hpp:297-339): the MRPD word,
// then one word per table while Mapped, stamping each table block read
// (compute_element_position, K_CacheDevice.hpp:316-327).  Returns the last
// word read: w = 1 maps the brick slot, 0 Unmapped, 2 Empty.  (The
// reference, after an Unmapped table entry with a coarser level pending,
// repeats the same lookup for ever, K_CacheManager.hpp:318-330; such an
// entry is Unmapped here, as an Unmapped MRPD entry is.)
__device__ inline uint32_t pt_walk(const DvrLaunch &p, const PtKey &k, uint32_t ts)
{
    uint32_t w = p.cells[k.top];
#pragma unroll
    for (int i = 0; i < GC_MAX_TABLES; ++i) {
        if ((uint32_t)i < p.pt.n && word_state(w) == 1) {
            const uint32_t b = word_slot(w);
            p.pt.usage[i][b] = ts;
            w = p.pt.words[i][(uint64_t)b * p.pt.b3[i] + k.loc[i]];
        }
    }
    return w;
}

// Byte offset of the transfer-function tables in a launch's dynamic LDS:
// after `head` bytes (level-switch table, if any) and the view table.
__host__ __device__ inline uint32_t tf_lds_base(uint32_t head, uint32_t n_views)
{
    return (head + n_views * (uint32_t)sizeof(DvrView) + 15u) & ~15u;
}

// Each distinct table of the launch is copied into LDS once (by the view that
// owns it), padded with one clamped entry on each side: dst[j] = tf[clamp(j-1)].
__device__ inline void load_tf_tables(const DvrView *vw, uint32_t n_views, float4 *tfl, int tid)
{
    for (uint32_t v = 0; v < n_views; ++v) {
        const uint32_t o = vw[v].tf_lds_off;
        if (!(o & DVR_TF_OWNER)) continue;
        const int N = (int)vw[v].tf_size;
        float4 *dst = tfl + (o & ~DVR_TF_OWNER);
        for (int i = tid; i < N + 2; i += LANES) dst[i] = vw[v].tf[i < 1 ? 0 : (i > N ? N - 1 : i - 1)];
    }
}

// MT: the launch's views use more than one LDS transfer-function table (each
// ray keeps its table's offset and size); otherwise there is one table, at
// offset 0, whose size the whole launch shares.
// BSC: the brick size when it is fixed at compile time (32 with covering 1,
// the bricks of the out-of-core configurations), else 0: constant row and
// plane strides turn the footprint's four loads into one address and three
// immediate offsets.
// PT: the page directory has table caches (the non-pipelined march with the
// walk through them; otherwise the MRPD maps bricks).
// HO: the hand-over of long rays to the group march is compiled in (used
// when p.yield_after != 0; the other variants carry none of its registers).
template <int BYTES, typename ACC, int TW, int TH, bool MT, int BSC = 0, bool PT = false, bool HO = false>
__global__ __launch_bounds__(LANES) __attribute__((amdgpu_waves_per_eu(sizeof(ACC) == 8 ? 1 : DVR_WAVES_PER_EU))) void k_dvr(const DvrLaunch p)
{
    static_assert(TW * TH == 64, "one wave per ray packet");
    __shared__ LevelLds lv[GC_MAX_LEVELS];
    __shared__ PtLds plv[PT ? GC_MAX_LEVELS : 1];
    // dynamic LDS: level-switch distances [nb_lods][LANES], the views, then
    // the launch's transfer functions (padded copies), when they fit
    extern __shared__ __align__(16) uint8_t dyn_lds[];
    const uint32_t nl = p.nb_lods;
    const uint32_t n_views = p.views ? p.n_views : 1u;
    float *tsw = reinterpret_cast<float *>(dyn_lds);
    DvrView *vw = reinterpret_cast<DvrView *>(dyn_lds + (size_t)nl * LANES * sizeof(float));
    float4 *tfl = reinterpret_cast<float4 *>(dyn_lds + tf_lds_base(nl * LANES * (uint32_t)sizeof(float), n_views));
    const int tid = threadIdx.x;
    const gc_geometry *g = p.geom;
    if (tid < (int)nl) {
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
        L.esq = p.esq[tid];
        if constexpr (PT) pt_lds_fill(plv[tid], g, tid);
    }
    if (p.views) {
        for (uint32_t i = tid; i < n_views; i += LANES) vw[i] = p.views[i];
    } else if (tid == 0) {
        vw[0] = p.view0;
    }
    __syncthreads();
    const bool tf_lds = BYTES != 4 && p.tf_lds_entries != 0;
    const double qskip = p.qskip;
    const bool exact = p.exact_alpha != 0;
    if (tf_lds) load_tf_tables(vw, n_views, tfl, tid);
    __syncthreads();
    const uint32_t tf_n1 = p.tf_lds_entries - 2u;   // the single table's size (!MT)
    // packets: first the ray list (64 entries each), then the views' tiles
    // wave-uniform values read from memory go to scalar registers, waited for
    // here rather than inside the march loop
    const uint32_t list_count = __builtin_amdgcn_readfirstlane(p.list ? *p.list_count : 0u);
    const uint32_t list_packets = (list_count + 63u) / 64u;
    const uint32_t n_packets = list_packets + p.n_packets;

    const uint32_t lane = tid & 63;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const uint32_t xcd = blockIdx.x & 7u;
    const uint32_t bs = BSC ? (uint32_t)BSC : g->brick_size, cov = BSC ? 1u : g->covering,
                   inner = BSC ? (uint32_t)BSC - 2u : g->inner;
    const uint32_t MX = g->mrpd_dim[0], MY = g->mrpd_dim[1];
    const uint64_t bbytes = (uint64_t)bs * bs * bs * BYTES;
    const uint32_t row = bs * BYTES, plane = bs * bs * BYTES;
    const uint8_t *pool = static_cast<const uint8_t *>(p.pool);
    // byte offset of texel (cov-1, cov-1, cov-1) inside a brick
    const uint32_t corner = cov >= 1 ? (cov - 1) * (plane + row + BYTES) : 0u;
    const bool fast_fetch = BYTES != 4 && cov >= 1;
    const float tstep = p.tstep;
    const uint32_t ts = __builtin_amdgcn_readfirstlane(p.ts_dev ? *p.ts_dev : p.ts);
    const int march = (int)p.march;

    // wave-uniform packet stream: the current packet and its next free slot
    uint32_t cur_pk = 0, cur_j = 64;
    bool exhausted = false;

    // per-lane ray state
    bool have = false;
    uint32_t v = 0, pix = 0;
    uint32_t tfw = 0;   // the ray's LDS table: entry offset | entries << 16
    float t = 0.f, tfar = 0.f, astop = 0.f;
    F3 pos = f3(0, 0, 0), dir = f3(0, 0, 0), dstep = f3(0, 0, 0);
    // current level: index, next switch distance and its parameters
    uint32_t lod = 0, lcoord = 0;
    float tnext = FINF, step = 0.f;
    F3 ior = f3(0, 0, 0), rdv = f3(0, 0, 0), ldf = f3(0, 0, 0), vvf = f3(0, 0, 0);
    // current brick: page-table cell, entry state, slot and base address
    uint32_t pcx = NONE, pcy = NONE, pcz = NONE, ew = 0, slot = 0, vkey = NONE;
    PtKey pk_prev{NONE, {NONE, NONE, NONE}};   // PT: the walk's last key
    uint32_t vsum = 0;   // box sum of the voxel vkey (non-pipelined march)
    const uint8_t *bbase = pool;
    Acc4<ACC> fc{0, 0, 0, 0};
    // FP32: exact opacity product of the blended samples and its stop
    // threshold (see opacity_factor)
    double Q = 1.0, thr = 0.0;
    uint32_t n_samples = 0, n_visits = 0;      // per lane
    uint32_t n_rays = 0, n_unmapped = 0;       // per wave
    // hand-over: the lane's samples before its current ray (in LDS, off the
    // register budget: read only in rounds where a hand-over may happen)
    __shared__ uint32_t ray_n0[HO ? LANES : 1];
    uint32_t rounds = 0;
    bool go_seen = false;

    auto set_level = [&](uint32_t l) {
        const LevelLds &L = lv[l];
        ior = f3(L.ior[0], L.ior[1], L.ior[2]);
        rdv = f3(L.rdv[0], L.rdv[1], L.rdv[2]);
        ldf = f3(L.ldf[0], L.ldf[1], L.ldf[2]);
        vvf = f3(L.vvf[0], L.vvf[1], L.vvf[2]);
        lcoord = L.lcoord;
        step = L.step;
        dstep = dir * step;
        tnext = l + 1 < nl ? tsw[(l + 1) * LANES + tid] : FINF;
    };

    // ------------------------------------------------- pipelined march state
    // Three samples of a ray are in flight: the front sample j+2 (its page
    // entry is being loaded), sample j+1 (its box footprint is being loaded)
    // and sample j (consumed: side effects, classification, compositing).
    // Every half-step issues exactly one page-entry load and four footprint
    // loads, so each value is waited for one half-step after it was asked
    // for, with the five newer loads still in flight.  Side effects (usage and
    // request stamps, counters, stored state) happen only when a sample is
    // consumed, in the reference's order, so the trace is the one-sample-at-
    // a-time trace.  The page cell doubles as the request cell
    // (realDivVirtual == 1).
    const bool piped = !PT && fast_fetch && p.pipelined != 0;
    // (the pipelined march, or the walk through table caches)
    const uint32_t budget = HO && (piped || PT) && p.yield_after != 0 ? p.yield_after : 0xffffffffu;
    // hb: stage B holds a sample (the front placed last half-step); hc: stage
    // C holds one; newray: the next sample placed is the ray's first
    bool hb = false, hc = false, newray = false;
    uint32_t cellB = 0, viB = 0, cellC = 0, slotC = 0, ewC = 0, viC = 0, sum_prev = 0;
    uint32_t lodB = 0, lodC = 0;   // levels of samples B and C
    float tC = 0.f;
    // position of sample C, kept (in LDS, off the register budget) only when
    // a view of the launch asks for exact continuation or the launch hands
    // rays over
    __shared__ F3 pos_c[LANES];
    const bool keep_pos = p.exact_resume != 0 || budget != 0xffffffffu;
    bool freshB = false, freshC = false, reuseC = false;
    uint32_t eX = 0u, eY = 0u;   // cell-table words in flight
    Quad wX{0u, 0u, 0u, 0u}, wY = wX;

    // page cell and in-brick voxel of the front sample (get_normalized,
    // K_CacheManager.hpp:267-361, with realDivVirtual == 1)
    auto locate = [&](uint32_t &cell, uint32_t &vi) {
        locate_cell(pos, ior, ldf, vvf, lcoord, MX, MY, inner, bs, cell, vi);
    };
    // the progressive state and the pixel of a ray that ends or stops
    // (RayCasters.cu:138-147,189-197)
    auto store_state = [&](float tt, F3 pp) {
        const DvrView &V = vw[v];
        gp(V.tbuf)[pix] = tt;
        if (V.rpos) gp(reinterpret_cast<gf4 *>(V.rpos))[pix] = gf4{pp.x, pp.y, pp.z, tt};
        if constexpr (sizeof(ACC) == 8) {
            gp(reinterpret_cast<gd4 *>(V.cstate))[pix] = gd4{fc.x, fc.y, fc.z, fc.w};
        } else {
            gp(reinterpret_cast<gf4 *>(V.cstate))[pix] = gf4{fc.x, fc.y, fc.z, fc.w};
            if (V.qbuf) gp(V.qbuf)[pix] = Q;
        }
        gp(V.pixels)[pix] = pack(fc.x, fc.y, fc.z, fc.w);
    };
    // Classification and front-to-back blending of one mapped sample
    // (RayCasters.cu:161-178); composite_if blends only where `on`.
    // Integer voxels are classified from their box sum (exact texture
    // coordinate), float voxels from the sample value s.
    // At FP32 the skip decision comes from the exact opacity factor and the
    // blended sample's factor joins Q.
    auto classify_any = [&](uint32_t sum, ACC s, double sd, uint32_t lx, bool &cskip, double &y) -> Acc4<ACC> {
        const LevelLds &L = lv[lx];
        const float kx = L.kexp;
        double x = 1.0;
        Acc4<ACC> cc;
        if constexpr (BYTES != 4) {
            cc = tf_lds ? (MT ? classify_sum_padded<BYTES, ACC>(sum, tfl + (tfw & 0xffffu), tfw >> 16, kx, cskip, x)
                              : classify_sum_padded<BYTES, ACC>(sum, tfl, tf_n1, kx, cskip, x))
                        : classify_sum<BYTES, ACC>(sum, gp(reinterpret_cast<const gf4 *>(vw[v].tf)), vw[v].tf_size, kx,
                                                   cskip, x);
        } else {
            if constexpr (sizeof(ACC) == 8)
                cc = classify<ACC>(s, gp(reinterpret_cast<const gf4 *>(vw[v].tf)), vw[v].tf_size, kx, cskip);
            else
                cc = classify_float_x(s, sd, gp(reinterpret_cast<const gf4 *>(vw[v].tf)), vw[v].tf_size, kx, cskip, x);
        }
        if constexpr (sizeof(ACC) == 4) {
            if (exact) {
                y = opacity_factor(x, L.esq);
                cskip = !(y < qskip);
            }
        }
        return cc;
    };
    auto composite_if = [&](bool on, uint32_t sum, uint32_t lx) {
        bool cskip;
        double y = 1.0;
        const Acc4<ACC> cc = classify_any(sum, (ACC)0, 0.0, lx, cskip, y);
        const bool b = on && !cskip;
        if constexpr (sizeof(ACC) == 4) Q = b ? Q * y : Q;
        const ACC wgt = (ACC)1 - fc.w;
        Acc4<ACC> n;
        if constexpr (sizeof(ACC) == 8) {
            n = Acc4<ACC>{fc.x + cc.x * wgt, fc.y + cc.y * wgt, fc.z + cc.z * wgt, fc.w + cc.w * wgt};
        } else {
            n = Acc4<ACC>{fmaf(cc.x, wgt, fc.x), fmaf(cc.y, wgt, fc.y), fmaf(cc.z, wgt, fc.z), fmaf(cc.w, wgt, fc.w)};
        }
        fc.x = b ? n.x : fc.x;
        fc.y = b ? n.y : fc.y;
        fc.z = b ? n.z : fc.z;
        fc.w = b ? n.w : fc.w;
    };
    auto composite = [&](uint32_t sum, ACC s, double sd, uint32_t lx) {
        bool cskip;
        double y = 1.0;
        const Acc4<ACC> cc = classify_any(sum, s, sd, lx, cskip, y);
        if (!cskip) {
            if constexpr (sizeof(ACC) == 4) Q *= y;
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
    };

    for (;;) {
        // ------------------------------------------------------------ refill
        const uint64_t idle = __ballot(!have);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (n_idle && !exhausted && (n_idle >= p.refill_min || n_idle == 64)) {
            const uint32_t avail = 64 - cur_j;
            uint32_t next_pk = NONE;
            if (n_idle > avail) {
                uint32_t got = NONE;
                if (lane == 0) got = fetch_packet(p, n_packets, xcd);
                next_pk = __shfl(got, 0, 64);
                if (next_pk == NONE) exhausted = true;
            }
            uint32_t my_pk = NONE, my_j = 0;
            if (!have) {
                const uint32_t rank = (uint32_t)__popcll(idle & lt_mask);
                if (rank < avail) { my_pk = cur_pk; my_j = cur_j + rank; }
                else if (next_pk != NONE && rank - avail < 64) { my_pk = next_pk; my_j = rank - avail; }
            }
            if (n_idle >= avail) {
                if (next_pk != NONE) { cur_pk = next_pk; cur_j = n_idle - avail; }
                else cur_j = 64;
            } else {
                cur_j += n_idle;
            }
            bool counted = false;
            if (my_pk != NONE) {
                // resolve the slot to (view, pixel)
                bool valid = false;
                uint32_t x = 0, y = 0;
                if (my_pk < list_packets) {
                    const uint64_t idx = (uint64_t)my_pk * 64 + my_j;
                    if (idx < list_count) {
                        const uint32_t e = p.list[idx];
                        v = e >> DVR_PIXEL_BITS;
                        const uint32_t px = e & ((1u << DVR_PIXEL_BITS) - 1u);
                        x = px % vw[v].width;
                        y = px / vw[v].width;
                        valid = true;
                    }
                } else {
                    const uint32_t tp = my_pk - list_packets;
                    uint32_t vv = 0;
                    while (vv + 1 < n_views && vw[vv + 1].packet_off <= tp) ++vv;
                    v = vv;
                    const uint32_t tt = tp - vw[v].packet_off;
                    const uint32_t pr = vw[v].tile_pair, ti = tt >> pr, half = tt & pr;
                    x = (ti % vw[v].tiles_x) * TW + my_j % TW;
                    y = (ti / vw[v].tiles_x) * (TH << pr) + half * TH + my_j / TW;
                    valid = x < vw[v].width && y < vw[v].height;
                }
                if (valid) {
                    // ray set-up (RayCasters.cu:57-120)
                    const DvrView &V = vw[v];
                    const uint32_t W = V.width;
                    if constexpr (MT) tfw = (V.tf_lds_off & ~DVR_TF_OWNER) | (V.tf_size << 16);
                    const RaySetup R = setup_ray(V, x, y);
                    const F3 origin = R.origin, dxn = R.dxn, dyn = R.dyn, dleft = R.dleft, dup = R.dup;
                    dir = R.dir;
                    float tnear = R.tnear;
                    tfar = R.tfar;
                    const bool hit = R.hit;
                    pix = y * W + x;
                    const float tb = gp(V.tbuf)[pix];
                    tnear = fmaxf(tnear, tb);
                    if constexpr (sizeof(ACC) == 8) {
                        const gd4 c = gp(reinterpret_cast<const gd4 *>(V.cstate))[pix];
                        fc = Acc4<ACC>{c.x, c.y, c.z, c.w};
                    } else {
                        const gf4 c = gp(reinterpret_cast<const gf4 *>(V.cstate))[pix];
                        fc = Acc4<ACC>{c.x, c.y, c.z, c.w};
                    }
                    counted = true;
                    if (!hit) {
                        fc.x = fc.y = fc.z = (ACC)0;
                        gp(V.pixels)[pix] = pack(fc.x, fc.y, fc.z, fc.w);
                    } else {
                        if (tnear < 0.f) tnear = 0.f;
                        t = tnear;
                        pos = origin + dir * tnear;
                        if (V.rpos && tb > 0.f && tb == tnear) {
                            // exact continuation of a stopped ray
                            const gf4 q = gp(reinterpret_cast<const gf4 *>(V.rpos))[pix];
                            pos = f3(q.x, q.y, q.z);
                        }
                        const float epx = fabsf(dot(dir - dxn, dleft) * tfar);
                        const float epy = fabsf(dot(dir - dyn, dup) * tfar);
                        astop = V.alpha_stop;
                        if constexpr (sizeof(ACC) == 4) {
                            thr = V.thr;
                            Q = (tb > 0.f && V.qbuf) ? gp(V.qbuf)[pix] : 1.0;
                        }
                        // level-switch distances from t on; the level at t
                        // is the last one already reached (compute_LOD)
                        uint32_t l0 = 0;
                        for (uint32_t l = 1; l < nl; ++l) {
                            const float T = level_switch(t, tfar, epx, epy, spacing(dleft, lv[l].inv),
                                                         spacing(dup, lv[l].inv));
                            tsw[l * LANES + tid] = T;
                            if (T <= t && l0 == l - 1) l0 = l;
                            if (T == FINF) break;
                        }
                        lod = l0;
                        set_level(l0);
                        pcx = pcy = pcz = NONE;
                        if constexpr (PT) pk_prev.top = NONE;
                        vkey = NONE;
                        hb = hc = false;
                        newray = true;
                        have = true;
                        if constexpr (HO) ray_n0[tid] = n_samples;
                    }
                }
            }
            n_rays += (uint32_t)__popcll(__ballot(counted));
        }
        if (__ballot(have) == 0) {
            if (exhausted) break;
            continue;
        }

        // ------------------------------------------------------------- march
        bool stopped = false;   // stopped at an Unmapped brick in this round
        // hand-over allowed in this round (wave-uniform): at once, or once
        // the go flag of the workgroup's line is up (read every 8th round)
        if (budget != 0xffffffffu && p.yield_exited != 0 && !go_seen && (++rounds & 7u) == 0u)
            go_seen = __builtin_amdgcn_readfirstlane(__hip_atomic_load(p.go + (blockIdx.x & 63u) * 32u, __ATOMIC_RELAXED,
                                                                       __HIP_MEMORY_SCOPE_AGENT)) != 0;
        const bool handoff_now = budget != 0xffffffffu && (p.yield_exited == 0 || go_seen);
        if constexpr (BYTES != 4) {
            if (piped) {
                // One half-step: issue the page entry of the new front sample
                // into e_new and the footprint of the previous front sample
                // into w_new; consume the sample whose footprint is w_old.
                // Halves alternate the X/Y registers, so nothing in flight is
                // ever copied.
                auto half = [&](uint32_t &e_new, const uint32_t &e_old, Quad &w_new, const Quad &w_old) {
                    // The common path is branch-free (selects): most lanes
                    // place, fetch and consume a sample every half-step.
                    // A: place the front sample (the reference's advance by
                    // the step of the level of the sample before it)
                    const float t_b = t;
                    const F3 pos_b = pos;
                    const uint32_t lod_b = lod;
                    {
                        const float ta = t + step;
                        const F3 pa = pos + dstep;
                        t = hb ? ta : t;
                        pos = f3(hb ? pa.x : pos.x, hb ? pa.y : pos.y, hb ? pa.z : pos.z);
                    }
                    if (__ballot(t >= tnext)) {
                        while (t >= tnext) set_level(++lod);
                    }
                    uint32_t cellA, viA;
                    locate(cellA, viA);
                    const bool freshA = newray || cellA != cellB;
                    // A lane that stays in its brick needs no new entry: it
                    // loads the wave's first cell instead (one shared line)
                    // and B keeps the entry it already has.
                    const uint32_t cell_u = (uint32_t)__builtin_amdgcn_readfirstlane(cellA);
                    e_new = p.cells[freshA ? cellA : cell_u];
                    // B: footprint of the sample placed last time.  Every
                    // lane issues the loads (so the waits stay one half-step
                    // behind); lanes that need none (not mapped, or the same
                    // voxel as sample j, whose box sum is kept) load the
                    // wave's first footprint instead.
                    const uint32_t slot_e = word_slot(e_old);
                    const uint32_t ewB = freshB ? word_state(e_old) : ewC;
                    const uint32_t slotB = freshB ? slot_e : slotC;
                    const bool reuseB = !freshB && viB == viC;
                    const uint64_t fr = (uint64_t)slotB * (uint32_t)bbytes + corner + viB * (uint32_t)BYTES;
                    // (readfirstlane returns int: widen through uint32_t,
                    // never by sign extension, for offsets of 2 GiB and up)
                    const uint64_t fu =
                        ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(fr >> 32)) << 32) |
                        (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)fr);
                    const bool needB = hb && ewB == 1 && !reuseB;
                    w_new = box_load<BYTES>(pool + (needB ? fr : fu), row, plane);
                    // Empty brick: the ray advances by tstep from it
                    // (RayCasters.cu:154-159), not by the level step.  The
                    // front sample just placed is dropped and placed again,
                    // from here, in the next half-step.
                    const bool discard = hb && ewB == 2;
                    if (__ballot(discard)) {
                        if (discard) {
                            t = t_b + tstep;
                            pos = pos_b + dir * tstep;
                            if (lod != lod_b) {
                                lod = lod_b;
                                set_level(lod);
                            }
                        }
                    }
                    // C: consume sample j (RayCasters.cu:118-192)
                    bool live;
                    if constexpr (sizeof(ACC) == 4) live = hc && (exact ? Q > thr : fc.w < astop) && tC < tfar;
                    else live = hc && fc.w < (ACC)astop && tC < tfar;
                    // hand-over: a ray past its budget stops before the
                    // first sample of a mapped brick it enters, so the group
                    // march that continues it visits and stamps that brick
                    // as this march would have
                    bool yld = false;
                    if (handoff_now) yld = live && ewC == 1 && freshC && n_samples - ray_n0[tid] >= budget;
                    const bool unmapped = live && ewC == 0;
                    const bool cons = live && ewC == 1 && !yld;
                    n_visits += (live && freshC && !yld) ? 1u : 0u;
                    if (unmapped) {
                        // raise_brick_request, then stop (Unmapped)
                        p.request[cellC] = ts;
                        stopped = true;
                    }
                    if (cons && freshC) p.usage[slotC] = ts;
                    // every lane classifies (a footprint sum is always a
                    // valid sample value); only consumed samples blend
                    const uint32_t sum = reuseC ? sum_prev : box_add<BYTES>(w_old);
                    sum_prev = cons ? sum : sum_prev;
                    composite_if(cons, sum, lodC);
                    n_samples += cons ? 1u : 0u;
                    const bool end = (hc && !live) || unmapped || yld;
                    if (end) store_state(tC, keep_pos ? pos_c[tid] : pos_b);
                    const uint64_t ym = HO ? __ballot(yld) : 0ull;
                    if (HO && ym) {
                        // the wave's handed-over rays, with one atomic
                        const uint32_t first = (uint32_t)__builtin_ctzll(ym);
                        uint32_t base = 0;
                        if (lane == first) base = atomicAdd(p.yield_count, (uint32_t)__popcll(ym));
                        base = __shfl(base, (int)first, 64);
                        if (yld) {
                            const uint32_t slot_y = base + (uint32_t)__popcll(ym & lt_mask);
                            p.yield_list[slot_y] = (v << DVR_PIXEL_BITS) | pix;
                            const F3 pc = pos_c[tid];
                            gp(reinterpret_cast<gf4 *>(p.yield_pos))[slot_y] = gf4{pc.x, pc.y, pc.z, tC};
                        }
                    }
                    // w_old is waited for on every path here (the main path
                    // already has), so its registers are free when the next
                    // half issues loads into them
                    asm volatile("" ::"v"(w_old.a), "v"(w_old.b), "v"(w_old.c), "v"(w_old.d));
                    // shift: B's sample moves to C (C keeps the last sample
                    // when B had none, as the reference of "same brick")
                    cellC = hb ? cellB : cellC;
                    slotC = hb ? slotB : slotC;
                    ewC = hb ? ewB : ewC;
                    tC = hb ? t_b : tC;
                    if (keep_pos && hb) pos_c[tid] = pos_b;
                    lodC = hb ? lodB : lodC;
                    freshC = hb ? freshB : freshC;
                    viC = hb ? viB : viC;
                    reuseC = hb ? reuseB : reuseC;
                    hc = hb && !end;
                    cellB = discard ? cellB : cellA;
                    viB = discard ? viB : viA;
                    lodB = discard ? lodB : lod;
                    freshB = discard ? freshB : freshA;
                    newray = newray && discard;
                    have = have && !end;
                    hb = have && !discard;
                };
                // Every lane runs every half-step: a lane without a ray
                // places and loads harmlessly (its page cell and footprint
                // stay in range; all side effects need hb/hc), and the loop
                // exits on a wave-uniform test after whole pairs, so the
                // waits above hold on every path.
                for (int k = 0; k < march; k += 2) {
                    half(eX, eY, wX, wY);
                    half(eY, eX, wY, wX);
                    if (__ballot(have) == 0) break;
                }
            }
        }
        bool yielded = false;   // handed over (non-pipelined march through table caches)
        if (!piped && have) {
            bool done = false;
            for (int k = 0; k < march; ++k) {
                bool go;
                if constexpr (sizeof(ACC) == 4) go = exact ? Q > thr : fc.w < astop;
                else go = fc.w < (ACC)astop;
                if (!(go && t < tfar)) { done = true; break; }
                while (t >= tnext) set_level(++lod);
                F3 tp = pos * 0.495f + f3(0.495f, 0.495f, 0.495f);
                tp.z = 1.0f - tp.z;
                const F3 np = f3(clamp099(tp.x), clamp099(tp.y), clamp099(tp.z));
                const F3 pr = np * ior;
                const F3 pv = pr * rdv;   // exact when rdv == 1
                uint32_t bx = 0, by = 0, bz = 0;
                if constexpr (!PT) {
                    bx = (uint32_t)(pv.x * ldf.x);
                    by = (uint32_t)(pv.y * ldf.y);
                    bz = (uint32_t)(pv.z * ldf.z);
                    const uint32_t mcx = lcoord + bx;
                    if (mcx != pcx || by != pcy || bz != pcz) {
                        pcx = mcx; pcy = by; pcz = bz;
                        // the page directory holds fewer than 2^32 cells (checked at creation)
                        const uint32_t e = p.cells[mcx + MX * (by + MY * bz)];
                        ew = word_state(e);
                        ++n_visits;
                        slot = word_slot(e);
                        bbase = pool + (uint64_t)slot * bbytes + corner;
                        vkey = NONE;
                        if (ew == 1) p.usage[slot] = ts;
                    }
                } else {
                    // the walk's key; walk again when it changes
                    const PtKey k = pt_key(p.pt, plv[lod], pv, bx, by, bz);
                    if (!pt_same(k, pk_prev)) {
                        pk_prev = k;
                        const uint32_t e = pt_walk(p, k, ts);
                        ew = word_state(e);
                        // hand-over before the visit of a mapped brick: the
                        // group march walks, stamps and counts it again
                        // (the walk's stamps carry the same timestamp)
                        if constexpr (HO) {
                            if (handoff_now && ew == 1 && n_samples - ray_n0[tid] >= budget) {
                                yielded = true;
                                break;
                            }
                        }
                        ++n_visits;
                        slot = word_slot(e);
                        bbase = pool + (uint64_t)slot * bbytes + corner;
                        vkey = NONE;
                        if (ew == 1) p.usage[slot] = ts;
                    }
                }
                if (ew == 0) {
                    // raise_brick_request with the real-volume position
                    const uint32_t rx = lcoord + (uint32_t)(pr.x * ldf.x);
                    const uint32_t ry = (uint32_t)(pr.y * ldf.y);
                    const uint32_t rz = (uint32_t)(pr.z * ldf.z);
                    p.request[rx + (uint64_t)MX * (ry + (uint64_t)MY * rz)] = ts;
                    stopped = true;
                    break;
                }
                if (ew != 1) {
                    // Empty brick: advance by tstep (RayCasters.cu:154-159);
                    // a ray that passes tfar ends at the next check
                    t += tstep;
                    pos = pos + dir * tstep;
                    continue;
                }
                // voxel within the brick: v mod (bs-2c), via the brick index
                // already known (falls back when FP32 rounding put the voxel
                // index in a neighbouring brick)
                const uint32_t vx = (uint32_t)(pv.x * vvf.x);
                const uint32_t vy = (uint32_t)(pv.y * vvf.y);
                const uint32_t vz = (uint32_t)(pv.z * vvf.z);
                // 24-bit products: voxel and brick indices stay far below 2^24
                uint32_t lx = vx - __umul24(bx, inner), ly = vy - __umul24(by, inner), lz = vz - __umul24(bz, inner);
                if (__builtin_expect(lx >= inner || ly >= inner || lz >= inner, 0)) {
                    lx = vx % inner;
                    ly = vy % inner;
                    lz = vz % inner;
                }
                ACC s = 0;
                double sd = 0.0;
                uint32_t isum = 0;
                if (fast_fetch) {
                    const uint32_t vi = __umul24(__umul24(lz, bs) + ly, bs) + lx;   // voxel in the brick
                    const uint32_t key = vi + (lod << 24);
                    if (key != vkey) {
                        vkey = key;
                        vsum = box_sum<BYTES == 4 ? 1 : BYTES>(bbase + vi * BYTES, row, plane);
                    }
                    isum = vsum;
                } else {
                    uint32_t o[3];
                    slot_origin(*g, slot, o);
                    const int ex = (int)(cov + o[0] + lx), ey = (int)(cov + o[1] + ly), ez = (int)(cov + o[2] + lz);
                    float fsum = 0.f;
                    double dsum = 0.0;
                    for (int kk = 0; kk < 2; ++kk)
                        for (int j = 0; j < 2; ++j)
                            for (int i = 0; i < 2; ++i) {
                                const float tv = texel<BYTES>(pool, g, ex - 1 + i, ey - 1 + j, ez - 1 + kk);
                                fsum += tv;
                                dsum += tv;
                            }
                    if constexpr (BYTES == 4) {
                        sd = dsum / 8.0;
                        if constexpr (sizeof(ACC) == 8) s = sd;
                        else s = fsum / 8.0f;
                    } else {
                        isum = (uint32_t)dsum;   // a sum of 8 integer voxels, exact
                    }
                }
                composite(isum, s, sd, lod);
                ++n_samples;
                t += step;
                pos = pos + dstep;
            }
            if (done || stopped || yielded) {
                store_state(t, pos);
                have = false;
            }
        }
        if constexpr (HO && PT) {
            const uint64_t ym = __ballot(yielded);
            if (ym) {
                const uint32_t first = (uint32_t)__builtin_ctzll(ym);
                uint32_t base = 0;
                if (lane == first) base = atomicAdd(p.yield_count, (uint32_t)__popcll(ym));
                base = __shfl(base, (int)first, 64);
                if (yielded) {
                    const uint32_t slot_y = base + (uint32_t)__popcll(ym & lt_mask);
                    p.yield_list[slot_y] = (v << DVR_PIXEL_BITS) | pix;
                    gp(reinterpret_cast<gf4 *>(p.yield_pos))[slot_y] = gf4{pos.x, pos.y, pos.z, t};
                }
            }
        }
        // rays stopped at a miss: flagged for the next launch's ray list
        const uint64_t st = __ballot(stopped);
        n_unmapped += (uint32_t)__popcll(st);
        if (stopped) {
            uint8_t *const sf = vw[v].stop;
            if (sf) gp(sf)[pix] = 1;
        }
    }

    if (HO && p.exited && lane == 0) {
        // the wave that completes the count raises the go flags
        const uint32_t n = atomicAdd(p.exited, 1u) + 1u;
        if (n == p.yield_exited)
            for (uint32_t k = 0; k < 64; ++k)
                __hip_atomic_store(p.go + k * 32u, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (p.stats) {
        const uint32_t s0 = wave_sum(n_samples), s1 = wave_sum(n_visits), s2 = n_rays, s3 = n_unmapped;
        if (lane == 0) {
            if (s0) atomicAdd((unsigned long long *)&p.stats->mapped_samples, (unsigned long long)s0);
            if (s1) atomicAdd((unsigned long long *)&p.stats->brick_visits, (unsigned long long)s1);
            if (s2) atomicAdd((unsigned long long *)&p.stats->pixels, (unsigned long long)s2);
            if (s3) atomicAdd((unsigned long long *)&p.stats->unmapped_rays, (unsigned long long)s3);
        }
    }
}

// ------------------------------------------------------------------------
// Wave-per-ray march, for launches that resume a short list of rays (the
// rays an earlier launch stopped at a miss).  Such launches hold too few rays
// to fill the machine and last as long as their longest ray, which the
// lane-per-ray march walks one sample per step at memory latency.  Here one
// wave walks one ray, 64 samples per super-step:
//  * the 64 sample positions are placed sequentially, in the reference's
//    FP32 order (t += step, position += direction * step, level switches),
//    and lane k keeps sample k;
//  * every lane locates, loads and classifies its own sample in parallel;
//  * the samples are then consumed in order, front to back, by a scan that
//    applies the reference's termination tests before each sample; page
//    visits, usage and request stamps and counters cover exactly the
//    consumed samples, so the trace equals the one-sample-at-a-time trace.
// Same configurations as the pipelined march (u8/u16 voxels, covering >= 1,
// realDivVirtual == 1).  Empty bricks end a super-step (the reference steps
// over them by tstep, which moves every later sample).
template <typename ACC>
__device__ inline ACC lane_read(ACC x, uint32_t k)
{
    if constexpr (sizeof(ACC) == 8) {
        const uint64_t b = __builtin_bit_cast(uint64_t, x);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, (int)k);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), (int)k);
        return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
    } else {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), (int)k));
    }
}

// BSC as for k_dvr: 32 when the bricks are 32 with covering 1, else 0.
// PT: the page directory has table caches: each lane walks its sample's
// entries (reads only; the stamps of the table blocks go with the sample's
// consumption) and requests go to the brick-grid cell of the real-volume
// position (realDivVirtual may differ from 1).
template <int BYTES, typename ACC, int G, int BSC = 0, bool PT = false>
__global__ __launch_bounds__(LANES) __attribute__((amdgpu_waves_per_eu(sizeof(ACC) == 8 ? 1 : 4))) void k_dvr_ray(const DvrLaunch p)
{
    static_assert(G == 4 || G == 8 || G == 16 || G == 32 || G == 64, "4, 8, 16, 32 or 64 lanes per ray");
    constexpr uint32_t R = 64 / G;   // rays per wave
    __shared__ LevelLds lv[GC_MAX_LEVELS];
    __shared__ PtLds plv[PT ? GC_MAX_LEVELS : 1];
    __shared__ float tswv[WAVES * R][GC_MAX_LEVELS + 1];
    // dynamic LDS: the views, then the launch's transfer functions
    extern __shared__ __align__(16) uint8_t dyn_lds[];
    DvrView *vw = reinterpret_cast<DvrView *>(dyn_lds);
    const uint32_t nl = p.nb_lods;
    const int tid = threadIdx.x;
    const gc_geometry *g = p.geom;
    if (tid < (int)nl) {
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
        L.esq = p.esq[tid];
        if constexpr (PT) pt_lds_fill(plv[tid], g, tid);
    }
    const uint32_t n_views = p.views ? p.n_views : 1u;
    if (p.views) {
        for (uint32_t i = tid; i < n_views; i += LANES) vw[i] = p.views[i];
    } else if (tid == 0) {
        vw[0] = p.view0;
    }
    __syncthreads();
    float4 *tfl = reinterpret_cast<float4 *>(dyn_lds + tf_lds_base(0, n_views));
    const bool tf_lds = p.tf_lds_entries != 0;
    if (tf_lds) load_tf_tables(vw, n_views, tfl, tid);
    __syncthreads();

    const uint32_t list_count = __builtin_amdgcn_readfirstlane(*p.list_count);
    const double qskip = p.qskip;
    const bool exact = p.exact_alpha != 0;
    const uint32_t bs = BSC ? (uint32_t)BSC : g->brick_size, cov = BSC ? 1u : g->covering,
                   inner = BSC ? (uint32_t)BSC - 2u : g->inner;
    const uint32_t MX = g->mrpd_dim[0], MY = g->mrpd_dim[1];
    const uint32_t bbytes = bs * bs * bs * BYTES;
    const uint32_t row = bs * BYTES, plane = bs * bs * BYTES;
    const uint8_t *pool = static_cast<const uint8_t *>(p.pool);
    const uint32_t corner = (cov - 1) * (plane + row + BYTES);
    const float tstep = p.tstep;
    const uint32_t ts = __builtin_amdgcn_readfirstlane(p.ts_dev ? *p.ts_dev : p.ts);

    const uint32_t lane = tid & 63, gl = lane % G, gsh = lane - gl;
    float *tsw = tswv[((uint32_t)tid >> 6) * R + lane / G];
    const uint64_t glow = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    auto gbits = [&](uint64_t m) -> uint64_t { return G == 64 ? m : ((m >> gsh) & glow); };
    auto first_clear = [&](uint64_t m) -> uint32_t {   // index of the first 0 bit of the group's mask, or G
        const uint64_t z = ~m & glow;
        return z ? (uint32_t)__builtin_ctzll(z) : (uint32_t)G;
    };

    // the group's ray (every lane of the group holds the same values)
    bool active = false, exhausted = false;
    uint32_t e = 0, v = 0, pix = 0;
    Acc4<ACC> fc{0, 0, 0, 0};
    float t = 0.f, tfar = 0.f, astop = 0.f, step = 0.f, tnext = FINF;
    F3 pos = f3(0, 0, 0), dir = f3(0, 0, 0), dstep = f3(0, 0, 0);
    uint32_t lod = 0, cell_last = NONE;
    PtKey key_last{NONE, {NONE, NONE, NONE}};   // PT: the last consumed sample's walk key
    double Q = 1.0, thr = 0.0;   // FP32: exact opacity product and its stop threshold
    uint32_t n_samples = 0, n_visits = 0, n_rays = 0, n_unmapped = 0;   // on each group's first lane
    auto set_level = [&](uint32_t l) {
        step = lv[l].step;
        dstep = dir * step;
        tnext = l + 1 < nl ? tsw[l + 1] : FINF;
    };

    const uint64_t lt_mask = (1ull << lane) - 1ull;
    for (;;) {
        // a group without a ray takes the next one (the whole group
        // together); the wave's groups take consecutive list entries with
        // one atomic
        const uint64_t want = __ballot(!active && !exhausted && gl == 0);
        if (!active && !exhausted) {
            uint32_t base = 0;
            if (lane == (uint32_t)__builtin_ctzll(want)) base = atomicAdd(&p.queue[0], (uint32_t)__popcll(want));
            base = __shfl(base, __builtin_ctzll(want), 64);
            uint32_t r = base + (uint32_t)__popcll(want & lt_mask);
            r = __shfl(r, 0, G);
            if (r >= list_count) {
                exhausted = true;
            } else {
                e = p.list[r];
                v = e >> DVR_PIXEL_BITS;
                pix = e & ((1u << DVR_PIXEL_BITS) - 1u);
                const DvrView &V = vw[v];
                const RaySetup Rs = setup_ray(V, pix % V.width, pix / V.width);
                const float tb = gp(V.tbuf)[pix];
                float tnear = fmaxf(Rs.tnear, tb);
                if constexpr (sizeof(ACC) == 8) {
                    const gd4 c = gp(reinterpret_cast<const gd4 *>(V.cstate))[pix];
                    fc = Acc4<ACC>{c.x, c.y, c.z, c.w};
                } else {
                    const gf4 c = gp(reinterpret_cast<const gf4 *>(V.cstate))[pix];
                    fc = Acc4<ACC>{c.x, c.y, c.z, c.w};
                }
                if (gl == 0 && !p.continuation) ++n_rays;
                if (!Rs.hit) {
                    fc.x = fc.y = fc.z = (ACC)0;
                    if (gl == 0) gp(V.pixels)[pix] = pack(fc.x, fc.y, fc.z, fc.w);
                } else {
                    if (tnear < 0.f) tnear = 0.f;
                    dir = Rs.dir;
                    tfar = Rs.tfar;
                    astop = V.alpha_stop;
                    if constexpr (sizeof(ACC) == 4) {
                        thr = V.thr;
                        Q = (tb > 0.f && V.qbuf) ? gp(V.qbuf)[pix] : 1.0;
                    }
                    t = tnear;
                    pos = Rs.origin + dir * tnear;
                    if (p.continuation) {
                        // a ray handed over by the lane march: its exact position
                        const gf4 q = gp(reinterpret_cast<const gf4 *>(p.yield_pos))[r];
                        pos = f3(q.x, q.y, q.z);
                    } else if (V.rpos && tb > 0.f && tb == tnear) {
                        // exact continuation of a stopped ray
                        const gf4 q = gp(reinterpret_cast<const gf4 *>(V.rpos))[pix];
                        pos = f3(q.x, q.y, q.z);
                    }
                    // level-switch distances, one level per lane; the level
                    // at t is the last one already reached (compute_LOD,
                    // RayCasters.cu:551-567)
                    const float epx = fabsf(dot(dir - Rs.dxn, Rs.dleft) * tfar);
                    const float epy = fabsf(dot(dir - Rs.dyn, Rs.dup) * tfar);
                    for (uint32_t l = gl; l < nl; l += G)
                        tsw[l] = l == 0 ? FINF
                                        : level_switch(t, tfar, epx, epy, spacing(Rs.dleft, lv[l].inv),
                                                       spacing(Rs.dup, lv[l].inv));
                    __builtin_amdgcn_wave_barrier();
                    lod = 0;
                    for (uint32_t l = 1; l < nl && tsw[l] <= t; ++l) lod = l;
                    set_level(lod);
                    cell_last = NONE;
                    if constexpr (PT) key_last.top = NONE;
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0) {
            if (__ballot(!exhausted) == 0) break;
            continue;
        }

        // place samples 0..G-1 of this super-step; lane k of a group keeps
        // sample k of its ray (groups without a ray run along harmlessly)
        float my_t = t;
        F3 my_pos = pos;
        uint32_t my_lod = lod;
        for (uint32_t k = 0; k < (uint32_t)G; ++k) {
            const bool mine = gl == k;
            my_t = mine ? t : my_t;
            my_pos = f3(mine ? pos.x : my_pos.x, mine ? pos.y : my_pos.y, mine ? pos.z : my_pos.z);
            my_lod = mine ? lod : my_lod;
            t += step;
            pos = pos + dstep;
            if (__ballot(t >= tnext)) {
                while (t >= tnext) set_level(++lod);
            }
        }
        // locate, load and classify each lane's sample
        const LevelLds &L = lv[my_lod];
        uint32_t cell, vi, en;
        bool fresh;
        PtKey key{};
        uint32_t tblk[GC_MAX_TABLES] = {NONE, NONE, NONE};   // PT: table blocks the walk read
        if constexpr (!PT) {
            locate_cell(my_pos, f3(L.ior[0], L.ior[1], L.ior[2]), f3(L.ldf[0], L.ldf[1], L.ldf[2]),
                        f3(L.vvf[0], L.vvf[1], L.vvf[2]), L.lcoord, MX, MY, inner, bs, cell, vi);
            en = p.cells[cell];
            const uint32_t prev = __shfl_up(cell, 1u, G);
            fresh = cell != (gl == 0 ? cell_last : prev);
        } else {
            F3 tp = my_pos * 0.495f + f3(0.495f, 0.495f, 0.495f);
            tp.z = 1.0f - tp.z;
            const F3 np = f3(clamp099(tp.x), clamp099(tp.y), clamp099(tp.z));
            const F3 pr = np * f3(L.ior[0], L.ior[1], L.ior[2]);
            const F3 pv = pr * f3(L.rdv[0], L.rdv[1], L.rdv[2]);
            uint32_t bx = 0, by = 0, bz = 0;
            key = pt_key(p.pt, plv[my_lod], pv, bx, by, bz);
            // the request cell (raise_brick_request, K_CacheManager.hpp:365-396)
            cell = L.lcoord + (uint32_t)(pr.x * L.ldf[0]) +
                   MX * ((uint32_t)(pr.y * L.ldf[1]) + MY * (uint32_t)(pr.z * L.ldf[2]));
            const uint32_t vx = (uint32_t)(pv.x * L.vvf[0]), vy = (uint32_t)(pv.y * L.vvf[1]),
                           vz = (uint32_t)(pv.z * L.vvf[2]);
            uint32_t lx = vx - __umul24(bx, inner), ly = vy - __umul24(by, inner), lz = vz - __umul24(bz, inner);
            if (__builtin_expect(lx >= inner || ly >= inner || lz >= inner, 0)) {
                lx = vx % inner;
                ly = vy % inner;
                lz = vz % inner;
            }
            vi = __umul24(__umul24(lz, bs) + ly, bs) + lx;
            en = p.cells[key.top];
#pragma unroll
            for (int i = 0; i < GC_MAX_TABLES; ++i) {
                if ((uint32_t)i < p.pt.n && word_state(en) == 1) {
                    tblk[i] = word_slot(en);
                    en = p.pt.words[i][(uint64_t)tblk[i] * p.pt.b3[i] + key.loc[i]];
                }
            }
            PtKey prev;
            prev.top = __shfl_up(key.top, 1u, G);
#pragma unroll
            for (int i = 0; i < GC_MAX_TABLES; ++i) prev.loc[i] = __shfl_up(key.loc[i], 1u, G);
            fresh = !pt_same(key, gl == 0 ? key_last : prev);
        }
        const uint32_t ew = word_state(en);
        const uint32_t slot = word_slot(en);
        const uint64_t off = (active && ew == 1) ? (uint64_t)slot * bbytes + corner + vi * (uint32_t)BYTES : 0ull;
        const uint32_t sum = box_add<BYTES>(box_load<BYTES>(pool + off, row, plane));
        bool skip;
        double x = 1.0, y = 1.0;
        const DvrView &V = vw[v];
        const Acc4<ACC> cc =
            tf_lds ? classify_sum_padded<BYTES, ACC>(sum, tfl + (V.tf_lds_off & ~DVR_TF_OWNER), V.tf_size, L.kexp, skip, x)
                   : classify_sum<BYTES, ACC>(sum, gp(reinterpret_cast<const gf4 *>(V.tf)), V.tf_size, L.kexp, skip, x);
        if constexpr (sizeof(ACC) == 4) {
            // exact opacity decisions (opacity_factor)
            if (exact) {
                y = opacity_factor(x, L.esq);
                skip = !(y < qskip);
            }
        }
        const uint64_t m_fresh = gbits(__ballot(fresh)), m_skip = gbits(__ballot(skip));
        const uint32_t u = first_clear(gbits(__ballot(ew == 1)));   // first sample not mapped
        const uint32_t f = first_clear(gbits(__ballot(my_t < tfar)));   // first sample with t >= tfar
        // consume in order (RayCasters.cu:118-178); kend: first sample not
        // consumed (the one that ends the ray, the one not mapped, or G)
        uint32_t kend = G;
        bool done = false, scanning = active;
        for (uint32_t k = 0; k < (uint32_t)G; ++k) {
            const ACC cx = __shfl(cc.x, (int)k, G), cy = __shfl(cc.y, (int)k, G), cz = __shfl(cc.z, (int)k, G),
                      cw = __shfl(cc.w, (int)k, G);
            double yk = 1.0;
            if constexpr (sizeof(ACC) == 4) yk = __shfl(y, (int)k, G);
            bool go;
            if constexpr (sizeof(ACC) == 4) go = exact ? Q > thr : fc.w < astop;
            else go = fc.w < (ACC)astop;
            if (scanning) {
                if (k >= f || !go) {
                    done = true;
                    scanning = false;
                    kend = k;
                } else if (k >= u) {
                    scanning = false;
                    kend = k;
                } else if (!((m_skip >> k) & 1ull)) {
                    const ACC wgt = (ACC)1 - fc.w;
                    if constexpr (sizeof(ACC) == 8) {
                        const ACC tx = cx * wgt, ty = cy * wgt, tz = cz * wgt, tw = cw * wgt;
                        fc.x += tx; fc.y += ty; fc.z += tz; fc.w += tw;
                    } else {
                        fc.x = fmaf(cx, wgt, fc.x);
                        fc.y = fmaf(cy, wgt, fc.y);
                        fc.z = fmaf(cz, wgt, fc.z);
                        fc.w = fmaf(cw, wgt, fc.w);
                        Q *= yk;
                    }
                }
            }
            if ((k & 3u) == 3u && __ballot(scanning) == 0) break;
        }
        // the sample at kend, fetched with every lane active
        const uint32_t ks = kend < (uint32_t)G ? kend : (uint32_t)G - 1u;
        const float t_k = __shfl(my_t, (int)ks, G);
        const float px_k = __shfl(my_pos.x, (int)ks, G), py_k = __shfl(my_pos.y, (int)ks, G),
                    pz_k = __shfl(my_pos.z, (int)ks, G);
        const uint32_t lod_k = __shfl(my_lod, (int)ks, G), cell_k = __shfl(cell, (int)ks, G),
                       ew_k = __shfl(ew, (int)ks, G);
        PtKey key_k{};
        if constexpr (PT) {
            key_k.top = __shfl(key.top, (int)ks, G);
#pragma unroll
            for (int i = 0; i < GC_MAX_TABLES; ++i) key_k.loc[i] = __shfl(key.loc[i], (int)ks, G);
        }
        if (active) {
            // samples 0..kend-1 were consumed (all mapped)
            const uint64_t below = kend >= 64 ? ~0ull : ((1ull << kend) - 1ull);
            if (gl == 0) {
                n_samples += kend;
                n_visits += (uint32_t)__popcll(m_fresh & below);
            }
            if (gl < kend && fresh) p.usage[slot] = ts;
            if constexpr (PT) {
                // the walks of the consumed samples, and of the one at kend
                // when the ray reached it (Unmapped or Empty), stamp the
                // table blocks they read
                if ((gl < kend || (gl == kend && !done)) && fresh) {
#pragma unroll
                    for (int i = 0; i < GC_MAX_TABLES; ++i)
                        if (tblk[i] != NONE) p.pt.usage[i][tblk[i]] = ts;
                }
            }
            auto store_state = [&](float tt) {
                if (gl == 0) {
                    gp(V.tbuf)[pix] = tt;
                    if (V.rpos) gp(reinterpret_cast<gf4 *>(V.rpos))[pix] = gf4{px_k, py_k, pz_k, tt};
                    if constexpr (sizeof(ACC) == 8) {
                        gp(reinterpret_cast<gd4 *>(V.cstate))[pix] = gd4{fc.x, fc.y, fc.z, fc.w};
                    } else {
                        gp(reinterpret_cast<gf4 *>(V.cstate))[pix] = gf4{fc.x, fc.y, fc.z, fc.w};
                        if (V.qbuf) gp(V.qbuf)[pix] = Q;
                    }
                    gp(V.pixels)[pix] = pack(fc.x, fc.y, fc.z, fc.w);
                }
            };
            if (done) {
                store_state(t_k);
                active = false;
            } else if (kend == (uint32_t)G) {
                cell_last = cell_k;
                if constexpr (PT) key_last = key_k;
            } else {
                // sample kend is reached and not mapped
                if (gl == 0 && ((m_fresh >> kend) & 1ull)) ++n_visits;
                if (ew_k == 0) {
                    // raise_brick_request, then stop (Unmapped)
                    if (gl == kend) p.request[cell] = ts;
                    store_state(t_k);
                    if (gl == 0) {
                        ++n_unmapped;
                        if (V.stop) gp(V.stop)[pix] = 1;
                    }
                    active = false;
                } else {
                    // Empty brick: advance by tstep from it (RayCasters.cu:154-159)
                    cell_last = cell_k;
                    if constexpr (PT) key_last = key_k;
                    t = t_k + tstep;
                    pos = f3(px_k, py_k, pz_k) + dir * tstep;
                    lod = lod_k;
                    set_level(lod);
                    while (t >= tnext) set_level(++lod);
                }
            }
        }
    }
    if (p.stats) {
        const uint32_t s0 = wave_sum(n_samples), s1 = wave_sum(n_visits), s2 = wave_sum(n_rays), s3 = wave_sum(n_unmapped);
        if (lane == 0) {
            if (s0) atomicAdd((unsigned long long *)&p.stats->mapped_samples, (unsigned long long)s0);
            if (s1) atomicAdd((unsigned long long *)&p.stats->brick_visits, (unsigned long long)s1);
            if (s2) atomicAdd((unsigned long long *)&p.stats->pixels, (unsigned long long)s2);
            if (s3) atomicAdd((unsigned long long *)&p.stats->unmapped_rays, (unsigned long long)s3);
        }
    }
}

// ------------------------------------------------------------------------
// Maximum projection (RayCastMOPImpl, RayCasters.cu:201-349): the ray keeps
// the largest sample value, with the DVR march's stepping (level steps,
// tstep over Empty bricks, stop at the first Unmapped brick, no early
// termination), and classifies it once: pixel = TF(max), no opacity
// correction or blending.  The colour state's x holds the normalised maximum
// (colorStateBuffer[i].x).  One lane per ray, one wave per packet of 64 rays:
// first the launch's ray list, then the views' 8x8 tiles.
//
// Integer voxels keep the maximum box sum (exact; the normalised value
// sum / D is monotone in it and converts back exactly) and classify it at
// the exact texture coordinate of tf_coord; float voxels keep the maximum
// sample.

// TF(value) with linear filtering and clamp addressing, no opacity correction.
template <int BYTES, typename ACC, typename TFP>
__device__ inline Acc4<ACC> tf_lookup_sum(uint32_t sum, TFP TF, uint32_t N)
{
    constexpr uint32_t D = BYTES == 1 ? 2040u : 524280u;
    if constexpr (sizeof(ACC) == 8) {
        const double xf = ((double)sum / (double)D) * (double)N - 0.5;
        const double fl = floor(xf), a = xf - fl;
        int i0 = (int)fl, i1 = i0 + 1;
        i0 = i0 < 0 ? 0 : (i0 > (int)N - 1 ? (int)N - 1 : i0);
        i1 = i1 < 0 ? 0 : (i1 > (int)N - 1 ? (int)N - 1 : i1);
        const gf4 T0 = TF[i0], T1 = TF[i1];
        const double b = 1.0 - a;
        return Acc4<ACC>{b * (double)T0.x + a * (double)T1.x, b * (double)T0.y + a * (double)T1.y,
                         b * (double)T0.z + a * (double)T1.z, b * (double)T0.w + a * (double)T1.w};
    } else {
        uint32_t q, rem;
        tf_coord<BYTES, true>(sum, N, q, rem);
        const uint32_t i0 = q == 0 ? 0u : (q - 1u > N - 1u ? N - 1u : q - 1u);
        const uint32_t i1 = q > N - 1u ? N - 1u : q;
        const gf4 T0 = TF[i0], T1 = TF[i1];
        const float a = (float)rem * (1.0f / (float)(2u * D));
        return Acc4<ACC>{fmaf(a, T1.x - T0.x, T0.x), fmaf(a, T1.y - T0.y, T0.y), fmaf(a, T1.z - T0.z, T0.z),
                         fmaf(a, T1.w - T0.w, T0.w)};
    }
}
template <typename ACC, typename TFP>
__device__ inline Acc4<ACC> tf_lookup_value(ACC s, TFP TF, uint32_t N)
{
    const ACC xf = s * (ACC)N - (ACC)0.5;
    const ACC fl = floor(xf), a = xf - fl;
    int i0 = (int)fl, i1 = i0 + 1;
    i0 = i0 < 0 ? 0 : (i0 > (int)N - 1 ? (int)N - 1 : i0);
    i1 = i1 < 0 ? 0 : (i1 > (int)N - 1 ? (int)N - 1 : i1);
    const gf4 T0 = TF[i0], T1 = TF[i1];
    const ACC b = (ACC)1 - a;
    return Acc4<ACC>{b * (ACC)T0.x + a * (ACC)T1.x, b * (ACC)T0.y + a * (ACC)T1.y, b * (ACC)T0.z + a * (ACC)T1.z,
                     b * (ACC)T0.w + a * (ACC)T1.w};
}

template <int BYTES, typename ACC>
__global__ __launch_bounds__(256) void k_mop(const DvrLaunch p)
{
    __shared__ LevelLds lv[GC_MAX_LEVELS];
    __shared__ PtLds plv[GC_MAX_LEVELS];
    const int tid = threadIdx.x;
    const gc_geometry *g = p.geom;
    const uint32_t nl = p.nb_lods;
    if (tid < (int)nl) {
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
        if (p.pt.n) pt_lds_fill(plv[tid], g, tid);
    }
    __syncthreads();
    const uint32_t lane = tid & 63u;
    const uint32_t list_count = p.list ? *p.list_count : 0u;
    const uint32_t list_packets = (list_count + 63u) / 64u;
    const uint32_t pk = blockIdx.x * 4u + (uint32_t)(tid >> 6);
    const uint32_t n_views = p.views ? p.n_views : 1u;
    if (pk >= list_packets + p.n_packets) return;   // whole waves
    // the lane's ray: (view, pixel)
    bool valid = false;
    uint32_t v = 0, x = 0, y = 0;
    if (pk < list_packets) {
        const uint64_t idx = (uint64_t)pk * 64u + lane;
        if (idx < list_count) {
            const uint32_t e = p.list[idx];
            v = e >> DVR_PIXEL_BITS;
            const uint32_t px = e & ((1u << DVR_PIXEL_BITS) - 1u);
            const uint32_t W = p.views ? p.views[v].width : p.view0.width;
            x = px % W;
            y = px / W;
            valid = true;
        }
    } else {
        const uint32_t tp = pk - list_packets;
        if (p.views)
            while (v + 1 < n_views && p.views[v + 1].packet_off <= tp) ++v;
        const DvrView &V = p.views ? p.views[v] : p.view0;
        const uint32_t tt = tp - V.packet_off;
        x = (tt % V.tiles_x) * 8u + lane % 8u;
        y = (tt / V.tiles_x) * 8u + lane / 8u;
        valid = x < V.width && y < V.height;
    }
    uint32_t n_samples = 0, n_visits = 0;
    bool stopped = false;
    if (valid) {
        const DvrView &V = p.views ? p.views[v] : p.view0;
        const RaySetup R = setup_ray(V, x, y);
        const uint32_t pix = y * V.width + x;
        const float tb = gp(V.tbuf)[pix];
        float tnear = fmaxf(R.tnear, tb);
        // the stored maximum
        constexpr double D = BYTES == 1 ? 2040.0 : (BYTES == 2 ? 524280.0 : 1.0);
        ACC cs[4];
        if constexpr (sizeof(ACC) == 8) {
            const gd4 c = gp(reinterpret_cast<const gd4 *>(V.cstate))[pix];
            cs[0] = c.x; cs[1] = c.y; cs[2] = c.z; cs[3] = c.w;
        } else {
            const gf4 c = gp(reinterpret_cast<const gf4 *>(V.cstate))[pix];
            cs[0] = c.x; cs[1] = c.y; cs[2] = c.z; cs[3] = c.w;
        }
        uint32_t msum = 0;                 // integer voxels: maximum box sum
        ACC mval = cs[0];                  // float voxels: maximum sample
        double mvd = (double)cs[0];
        if constexpr (BYTES != 4) {
            const double r = rint((double)cs[0] * D);
            msum = r <= 0.0 ? 0u : (uint32_t)r;
        }
        if (!R.hit) {
            gp(V.pixels)[pix] = 0u;
        } else {
            const uint32_t bs = g->brick_size, cov = g->covering, inner = g->inner;
            const uint32_t MX = g->mrpd_dim[0], MY = g->mrpd_dim[1];
            const uint64_t bbytes = (uint64_t)bs * bs * bs * BYTES;
            const uint32_t row = bs * BYTES, plane = bs * bs * BYTES;
            const uint32_t corner = cov >= 1 ? (cov - 1) * (plane + row + BYTES) : 0u;
            const uint8_t *pool = static_cast<const uint8_t *>(p.pool);
            const uint32_t ts = p.ts_dev ? *p.ts_dev : p.ts;
            if (tnear < 0.f) tnear = 0.f;
            float t = tnear;
            const float tfar = R.tfar;
            F3 pos = R.origin + R.dir * tnear;
            if (V.rpos && tb > 0.f && tb == tnear) {
                const gf4 q = gp(reinterpret_cast<const gf4 *>(V.rpos))[pix];
                pos = f3(q.x, q.y, q.z);
            }
            const F3 dir = R.dir;
            const float epx = fabsf(dot(dir - R.dxn, R.dleft) * tfar);
            const float epy = fabsf(dot(dir - R.dyn, R.dup) * tfar);
            // level at t: the last one of the consecutive levels already
            // reached (compute_LOD); the next switch computed when needed
            uint32_t lod = 0;
            float tnext = FINF;
            for (uint32_t l = 1; l < nl; ++l) {
                const float T = level_switch(t, tfar, epx, epy, spacing(R.dleft, lv[l].inv), spacing(R.dup, lv[l].inv));
                if (T <= t) {
                    lod = l;
                } else {
                    tnext = T;
                    break;
                }
            }
            uint32_t pcell = NONE, ew = 0, slot = 0;
            PtKey pkp{NONE, {NONE, NONE, NONE}};
            for (;;) {
                if (!(t < tfar)) break;
                while (t >= tnext) {
                    ++lod;
                    tnext = lod + 1 < nl ? level_switch(t, tfar, epx, epy, spacing(R.dleft, lv[lod + 1].inv),
                                                        spacing(R.dup, lv[lod + 1].inv))
                                         : FINF;
                }
                const LevelLds &L = lv[lod];
                F3 tp = pos * 0.495f + f3(0.495f, 0.495f, 0.495f);
                tp.z = 1.0f - tp.z;
                const F3 np = f3(clamp099(tp.x), clamp099(tp.y), clamp099(tp.z));
                const F3 pr = np * f3(L.ior[0], L.ior[1], L.ior[2]);
                const F3 pv = pr * f3(L.rdv[0], L.rdv[1], L.rdv[2]);
                uint32_t bx = 0, by = 0, bz = 0;
                if (p.pt.n) {
                    const PtKey k = pt_key(p.pt, plv[lod], pv, bx, by, bz);
                    if (!pt_same(k, pkp)) {
                        pkp = k;
                        const uint32_t w = pt_walk(p, k, ts);
                        ew = word_state(w);
                        slot = word_slot(w);
                        ++n_visits;
                        if (ew == 1) p.usage[slot] = ts;
                    }
                } else {
                    bx = (uint32_t)(pv.x * L.ldf[0]);
                    by = (uint32_t)(pv.y * L.ldf[1]);
                    bz = (uint32_t)(pv.z * L.ldf[2]);
                    const uint32_t cell = L.lcoord + bx + MX * (by + MY * bz);
                    if (cell != pcell) {
                        pcell = cell;
                        const uint32_t w = p.cells[cell];
                        ew = word_state(w);
                        slot = word_slot(w);
                        ++n_visits;
                        if (ew == 1) p.usage[slot] = ts;
                    }
                }
                if (ew == 0) {
                    // raise_brick_request with the real-volume position, stop
                    const uint32_t rx = L.lcoord + (uint32_t)(pr.x * L.ldf[0]);
                    const uint32_t ry = (uint32_t)(pr.y * L.ldf[1]);
                    const uint32_t rz = (uint32_t)(pr.z * L.ldf[2]);
                    p.request[rx + (uint64_t)MX * (ry + (uint64_t)MY * rz)] = ts;
                    stopped = true;
                    break;
                }
                if (ew != 1) {
                    // Empty brick: tstep (RayCasters.cu:308-316)
                    t += p.tstep;
                    pos = pos + dir * p.tstep;
                    continue;
                }
                const uint32_t vx = (uint32_t)(pv.x * L.vvf[0]), vy = (uint32_t)(pv.y * L.vvf[1]),
                               vz = (uint32_t)(pv.z * L.vvf[2]);
                uint32_t lx = vx - bx * inner, ly = vy - by * inner, lz = vz - bz * inner;
                if (lx >= inner || ly >= inner || lz >= inner) {
                    lx = vx % inner;
                    ly = vy % inner;
                    lz = vz % inner;
                }
                if (BYTES != 4 && cov >= 1) {
                    const uint32_t vi = (lz * bs + ly) * bs + lx;
                    const uint32_t sum =
                        box_sum<BYTES == 4 ? 1 : BYTES>(pool + (uint64_t)slot * bbytes + corner + (uint64_t)vi * BYTES, row, plane);
                    msum = sum > msum ? sum : msum;
                } else {
                    uint32_t o[3];
                    slot_origin(*g, slot, o);
                    const int ex = (int)(cov + o[0] + lx), ey = (int)(cov + o[1] + ly), ez = (int)(cov + o[2] + lz);
                    float fsum = 0.f;
                    double dsum = 0.0;
                    for (int kk = 0; kk < 2; ++kk)
                        for (int j = 0; j < 2; ++j)
                            for (int i = 0; i < 2; ++i) {
                                const float tv = texel<BYTES>(pool, g, ex - 1 + i, ey - 1 + j, ez - 1 + kk);
                                fsum += tv;
                                dsum += tv;
                            }
                    if constexpr (BYTES == 4) {
                        const double sd = dsum / 8.0;
                        const ACC sv = sizeof(ACC) == 8 ? (ACC)sd : (ACC)(fsum / 8.0f);
                        if (sd > mvd) {
                            mvd = sd;
                            mval = sv;
                        }
                    } else {
                        const uint32_t sum = (uint32_t)dsum;
                        msum = sum > msum ? sum : msum;
                    }
                }
                ++n_samples;
                t += L.step;
                pos = pos + dir * L.step;
            }
            // state (RayCasters.cu:288-329) and pixel = TF(max)
            ACC mx;
            Acc4<ACC> c;
            const auto TF = gp(reinterpret_cast<const gf4 *>(V.tf));
            if constexpr (BYTES != 4) {
                mx = (ACC)((double)msum / D);
                c = tf_lookup_sum<BYTES, ACC>(msum, TF, V.tf_size);
            } else {
                mx = mval;
                c = tf_lookup_value<ACC>(mval, TF, V.tf_size);
            }
            gp(V.tbuf)[pix] = t;
            if (V.rpos) gp(reinterpret_cast<gf4 *>(V.rpos))[pix] = gf4{pos.x, pos.y, pos.z, t};
            if constexpr (sizeof(ACC) == 8) gp(reinterpret_cast<gd4 *>(V.cstate))[pix] = gd4{mx, cs[1], cs[2], cs[3]};
            else gp(reinterpret_cast<gf4 *>(V.cstate))[pix] = gf4{mx, cs[1], cs[2], cs[3]};
            gp(V.pixels)[pix] = pack(c.x, c.y, c.z, c.w);
            if (stopped && V.stop) gp(V.stop)[pix] = 1;
        }
    }
    if (p.stats) {
        const uint32_t s0 = wave_sum(n_samples), s1 = wave_sum(n_visits);
        const uint32_t s2 = (uint32_t)__popcll(__ballot(valid)), s3 = (uint32_t)__popcll(__ballot(stopped));
        if (lane == 0) {
            if (s0) atomicAdd((unsigned long long *)&p.stats->mapped_samples, (unsigned long long)s0);
            if (s1) atomicAdd((unsigned long long *)&p.stats->brick_visits, (unsigned long long)s1);
            if (s2) atomicAdd((unsigned long long *)&p.stats->pixels, (unsigned long long)s2);
            if (s3) atomicAdd((unsigned long long *)&p.stats->unmapped_rays, (unsigned long long)s3);
        }
    }
}

template <int BYTES, typename ACC, bool HO>
const void *kernel_of(uint32_t tile, bool mt, bool bs32, bool pt)
{
    // table caches: one variant per packet shape (transfer functions from
    // HBM when a launch has several)
    if (pt) {
        if (BYTES != 4 && sizeof(ACC) == 4 && (tile & 3) == 0 && bs32)
            return reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 8, 8, false, 32, true, HO>);
        switch (tile & 3) {
        case 0: return reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 8, 8, false, 0, true, HO>);
        case 1: return reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 16, 4, false, 0, true, HO>);
        case 2: return reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 32, 2, false, 0, true, HO>);
        default: return reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 64, 1, false, 0, true, HO>);
        }
    }
    // 8x8 tiles of u8/u16 bricks of 32 with covering 1 at FP32: constant strides
    if (BYTES != 4 && sizeof(ACC) == 4 && (tile & 3) == 0 && bs32)
        return mt ? reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 8, 8, true, 32, false, HO>)
                  : reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 8, 8, false, 32, false, HO>);
    // float volumes never use LDS tables: one variant
    if (BYTES == 4 || !mt) {
        switch (tile & 3) {
        case 0: return reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 8, 8, false, 0, false, HO>);
        case 1: return reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 16, 4, false, 0, false, HO>);
        case 2: return reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 32, 2, false, 0, false, HO>);
        default: return reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 64, 1, false, 0, false, HO>);
        }
    }
    switch (tile & 3) {
    case 0: return reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 8, 8, true, 0, false, HO>);
    case 1: return reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 16, 4, true, 0, false, HO>);
    case 2: return reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 32, 2, true, 0, false, HO>);
    default: return reinterpret_cast<const void *>(&k_dvr<BYTES, ACC, 64, 1, true, 0, false, HO>);
    }
}

// ho: the launch hands rays over (the FP32 u8/u16 marches have variants with
// and without it; the FP64 ones, at 2 waves per SIMD with registers to
// spare, always carry it; float volumes never hand over)
const void *dvr_kernel(uint32_t bytes, uint32_t precision, uint32_t tile, bool mt, bool bs32, bool pt = false,
                       bool ho = false)
{
    const bool d = precision == 64;
    tile = packet_kind(tile);
    switch (bytes) {
    case 1: return d ? kernel_of<1, double, true>(tile, mt, bs32, pt)
                     : (ho ? kernel_of<1, float, true>(tile, mt, bs32, pt) : kernel_of<1, float, false>(tile, mt, bs32, pt));
    case 2: return d ? kernel_of<2, double, true>(tile, mt, bs32, pt)
                     : (ho ? kernel_of<2, float, true>(tile, mt, bs32, pt) : kernel_of<2, float, false>(tile, mt, bs32, pt));
    default: return d ? kernel_of<4, double, false>(tile, mt, bs32, pt) : kernel_of<4, float, false>(tile, mt, bs32, pt);
    }
}

__global__ void k_fill32(uint32_t *p, uint32_t n, uint32_t value)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

} // namespace

uint32_t dvr_lds_bytes(uint32_t nb_lods, uint32_t n_views, uint32_t tf_entries)
{
    return tf_lds_base(nb_lods * LANES * (uint32_t)sizeof(float), std::max(1u, n_views)) + tf_entries * 16u;
}

uint32_t dvr_assign_tf_lds(DvrView *views, uint32_t n, uint32_t bytes, uint32_t *tables)
{
    *tables = 0;
    // distinct tables in order of first use
    const float4 *tabs[DVR_TF_LDS_TABLES];
    uint32_t size[DVR_TF_LDS_TABLES], off[DVR_TF_LDS_TABLES];
    uint32_t nt = 0, entries = 0;
    bool fits = bytes != 4;
    for (uint32_t i = 0; i < n && fits; ++i) {
        DvrView &V = views[i];
        V.tf_lds_off = 0;
        if (V.width == 0 || !V.tf) continue;   // unused slot
        uint32_t k = 0;
        while (k < nt && !(tabs[k] == V.tf && size[k] == V.tf_size)) ++k;
        if (k == nt) {
            if (nt == DVR_TF_LDS_TABLES || V.tf_size == 0 || V.tf_size > DVR_TF_LDS_MAX) {
                fits = false;
                break;
            }
            tabs[nt] = V.tf;
            size[nt] = V.tf_size;
            off[nt] = entries;
            entries += V.tf_size + 2;
            ++nt;
            V.tf_lds_off = off[k] | DVR_TF_OWNER;
        } else {
            V.tf_lds_off = off[k];
        }
    }
    if (!fits) {
        for (uint32_t i = 0; i < n; ++i) views[i].tf_lds_off = 0;
        return 0;
    }
    *tables = nt;
    return entries;
}

uint32_t dvr_grid(uint32_t bytes, uint32_t precision, uint32_t tile, uint32_t n_packets, uint32_t lds_bytes,
                  uint32_t tf_tables)
{
    // enough resident workgroups to fill every CU, never more than the packets
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dvr_kernel(bytes, precision, tile, tf_tables > 1, false),
                                                     LANES, lds_bytes) != hipSuccess ||
        per_cu <= 0)
        per_cu = 2;
    const uint32_t want = (uint32_t)cus * (uint32_t)per_cu;
    const uint32_t need = (n_packets + WAVES - 1) / WAVES;
    return std::max(1u, std::min(want, need));
}

hipError_t launch_dvr(const DvrLaunch &p, uint32_t bytes, uint32_t precision, uint32_t tile, uint32_t grid,
                      hipStream_t stream)
{
    if ((!p.list && p.n_packets == 0) || grid == 0) return hipSuccess;
    if (p.n_packets && !p.views && p.n_views != 1) return hipErrorInvalidValue;
    if (bytes != 1 && bytes != 2 && bytes != 4) return hipErrorInvalidValue;
    if (p.nb_lods == 0 || p.nb_lods > GC_MAX_LEVELS) return hipErrorInvalidValue;
    if (p.views && (p.n_views == 0 || p.n_views > DVR_MAX_VIEWS)) return hipErrorInvalidValue;
    if (p.list && !p.list_count) return hipErrorInvalidValue;
    if (p.march < 1 || p.march > 8) return hipErrorInvalidValue;
    if (p.yield_after && (!p.yield_list || !p.yield_count || !p.yield_pos || bytes == 4)) return hipErrorInvalidValue;
    if (p.yield_exited && (!p.exited || !p.go)) return hipErrorInvalidValue;
    if (p.pt.n) {
        if (p.pipelined) return hipErrorInvalidValue;
        // the table-cache variants keep one LDS transfer function at most
        DvrLaunch q = p;
        if (q.tf_lds_tables > 1) q.tf_lds_entries = q.tf_lds_tables = 0;
        void *args[] = {&q};
        return hipLaunchKernel(dvr_kernel(bytes, precision, tile, false, q.brick_size == 32 && q.covering == 1, true,
                                          q.yield_after != 0),
                               dim3(grid), dim3(LANES), args,
                               dvr_lds_bytes(q.nb_lods, q.views ? q.n_views : 1, q.tf_lds_entries), stream);
    }
    void *args[] = {const_cast<DvrLaunch *>(&p)};
    const bool bs32 = p.brick_size == 32 && p.covering == 1;
    return hipLaunchKernel(dvr_kernel(bytes, precision, tile, p.tf_lds_tables > 1, bs32, false, p.yield_after != 0),
                           dim3(grid), dim3(LANES), args,
                           dvr_lds_bytes(p.nb_lods, p.views ? p.n_views : 1, p.tf_lds_entries), stream);
}

uint32_t dvr_ray_lds_bytes(uint32_t n_views, uint32_t tf_entries)
{
    return tf_lds_base(0, std::max(1u, n_views)) + tf_entries * 16u;
}

bool dvr_ray_supported(const DvrLaunch &p, uint32_t bytes)
{
    // the pipelined march's configurations, or table caches with covering >= 1
    return (p.pt.n ? p.covering >= 1 : p.pipelined != 0) && (bytes == 1 || bytes == 2);
}

// Lanes per ray: a whole wave for the shortest lists (lowest latency per
// ray), 16 lanes (4 rays per wave) otherwise.
// Lanes per ray for a list of `rays`: the more rays, the fewer lanes each,
// so that longer lists keep the machine full with shorter per-ray chains.
uint32_t dvr_ray_group(uint32_t rays)
{
    if (rays <= 256u) return 64u;
    if (rays <= RAY_G16_MAX) return 16u;
    if (rays <= RAY_G8_MAX) return 8u;
    return 4u;
}

template <int BYTES, typename ACC>
static const void *ray_kernel_of(uint32_t group, bool bs32, bool pt)
{
    if (pt) {
        if (sizeof(ACC) == 4 && bs32) {
            switch (group) {
            case 16: return reinterpret_cast<const void *>(&k_dvr_ray<BYTES, ACC, 16, 32, true>);
            case 64: return reinterpret_cast<const void *>(&k_dvr_ray<BYTES, ACC, 64, 32, true>);
            default: return nullptr;
            }
        }
        switch (group) {
        case 16: return reinterpret_cast<const void *>(&k_dvr_ray<BYTES, ACC, 16, 0, true>);
        case 64: return reinterpret_cast<const void *>(&k_dvr_ray<BYTES, ACC, 64, 0, true>);
        default: return nullptr;
        }
    }
    if (sizeof(ACC) == 4 && bs32) {
        switch (group) {
        case 4: return reinterpret_cast<const void *>(&k_dvr_ray<BYTES, ACC, 4, 32>);
        case 8: return reinterpret_cast<const void *>(&k_dvr_ray<BYTES, ACC, 8, 32>);
        case 16: return reinterpret_cast<const void *>(&k_dvr_ray<BYTES, ACC, 16, 32>);
        case 64: return reinterpret_cast<const void *>(&k_dvr_ray<BYTES, ACC, 64, 32>);
        default: return nullptr;
        }
    }
    switch (group) {
    case 4: return reinterpret_cast<const void *>(&k_dvr_ray<BYTES, ACC, 4>);
    case 8: return reinterpret_cast<const void *>(&k_dvr_ray<BYTES, ACC, 8>);
    case 16: return reinterpret_cast<const void *>(&k_dvr_ray<BYTES, ACC, 16>);
    case 64: return reinterpret_cast<const void *>(&k_dvr_ray<BYTES, ACC, 64>);
    default: return nullptr;
    }
}

static const void *ray_kernel(uint32_t bytes, uint32_t precision, uint32_t group, bool bs32 = false, bool pt = false)
{
    const bool d = precision == 64;
    if (bytes == 1) return d ? ray_kernel_of<1, double>(group, bs32, pt) : ray_kernel_of<1, float>(group, bs32, pt);
    return d ? ray_kernel_of<2, double>(group, bs32, pt) : ray_kernel_of<2, float>(group, bs32, pt);
}

uint32_t dvr_ray_grid(uint32_t bytes, uint32_t precision, uint32_t rays, uint32_t lds_bytes, uint32_t group)
{
    // resident workgroups of 4 waves, never more groups than rays
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    if (group == 0) group = dvr_ray_group(rays);
    const void *k = ray_kernel(bytes, precision, group);
    if (!k) return 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, LANES, lds_bytes) != hipSuccess || per_cu <= 0)
        per_cu = 2;
    const uint32_t want = (uint32_t)cus * (uint32_t)per_cu;
    const uint64_t rays_per_block = (uint64_t)WAVES * (64u / group);
    const uint32_t need = (uint32_t)std::min<uint64_t>(((uint64_t)rays + rays_per_block - 1) / rays_per_block, want);
    return std::max(1u, need);
}

hipError_t launch_mop(const DvrLaunch &p, uint32_t bytes, uint32_t precision, uint32_t list_rays, hipStream_t stream)
{
    if (bytes != 1 && bytes != 2 && bytes != 4) return hipErrorInvalidValue;
    if (p.nb_lods == 0 || p.nb_lods > GC_MAX_LEVELS) return hipErrorInvalidValue;
    if (p.views && (p.n_views == 0 || p.n_views > DVR_MAX_VIEWS)) return hipErrorInvalidValue;
    if (p.list && !p.list_count) return hipErrorInvalidValue;
    const uint64_t packets = (uint64_t)(p.list ? (list_rays + 63u) / 64u : 0u) + p.n_packets;
    if (packets == 0) return hipSuccess;
    const void *k = nullptr;
    if (precision == 64) k = bytes == 1 ? reinterpret_cast<const void *>(&k_mop<1, double>)
                             : bytes == 2 ? reinterpret_cast<const void *>(&k_mop<2, double>)
                                          : reinterpret_cast<const void *>(&k_mop<4, double>);
    else k = bytes == 1 ? reinterpret_cast<const void *>(&k_mop<1, float>)
             : bytes == 2 ? reinterpret_cast<const void *>(&k_mop<2, float>)
                          : reinterpret_cast<const void *>(&k_mop<4, float>);
    void *args[] = {const_cast<DvrLaunch *>(&p)};
    return hipLaunchKernel(k, dim3((uint32_t)((packets + 3u) / 4u)), dim3(256), args, 0, stream);
}
hipError_t launch_dvr_ray(const DvrLaunch &p, uint32_t bytes, uint32_t precision, uint32_t rays, uint32_t grid,
                          hipStream_t stream, uint32_t group)
{
    if (!p.list || !p.list_count || p.n_packets != 0 || grid == 0) return hipErrorInvalidValue;
    if (p.continuation && !p.yield_pos) return hipErrorInvalidValue;
    if (!dvr_ray_supported(p, bytes)) return hipErrorInvalidValue;
    if (p.nb_lods == 0 || p.nb_lods > GC_MAX_LEVELS) return hipErrorInvalidValue;
    if (p.views && (p.n_views == 0 || p.n_views > DVR_MAX_VIEWS)) return hipErrorInvalidValue;
    if (group == 0) group = dvr_ray_group(rays);
    if (p.pt.n != 0 && group < 16u) group = 16u;   // no narrower table-walk variants
    const void *k = ray_kernel(bytes, precision, group, p.brick_size == 32 && p.covering == 1, p.pt.n != 0);
    if (!k) return hipErrorInvalidValue;
    void *args[] = {const_cast<DvrLaunch *>(&p)};
    return hipLaunchKernel(k, dim3(grid), dim3(LANES), args, dvr_ray_lds_bytes(p.views ? p.n_views : 1, p.tf_lds_entries),
                           stream);
}

hipError_t launch_fill32(uint32_t *p, uint32_t n, uint32_t value, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    k_fill32<<<(n + 255) / 256, 256, 0, stream>>>(p, n, value);
    return hipGetLastError();
}

// Each workgroup gathers 16 consecutive 8x8 tiles of the views' stop flags
// (four per wave, in tile order, so consecutive workgroups append
// neighbouring tiles): a wave reads and clears its tiles' flags, the
// workgroup reserves room for all its rays with one atomic, and each tile's
// rays are written contiguously.  One atomic per workgroup rather than per
// tile keeps the list counter off the critical path.
constexpr uint32_t COLLECT_TILES_PER_WAVE = 4, COLLECT_TILES = 4 * COLLECT_TILES_PER_WAVE;

__device__ inline uint32_t collect_view(const CollectArgs &a, uint32_t tile)
{
    // the last view whose range starts at or before the tile (views without
    // tiles share the next view's start)
    if (!a.views) return 0;
    uint32_t lo = 0, hi = a.n_views;   // stop_tile_off[lo] <= tile < stop_tile_off[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.views[mid].stop_tile_off <= tile) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_collect_stopped(const CollectArgs a)
{
    __shared__ uint32_t cnt[COLLECT_TILES];
    __shared__ uint32_t vof[COLLECT_TILES];
    __shared__ uint32_t base_s;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint64_t masks[COLLECT_TILES_PER_WAVE];
    uint32_t pixs[COLLECT_TILES_PER_WAVE], views[COLLECT_TILES_PER_WAVE];
    for (uint32_t k = 0; k < COLLECT_TILES_PER_WAVE; ++k) {
        const uint32_t j = wave * COLLECT_TILES_PER_WAVE + k;
        const uint32_t tile = blockIdx.x * COLLECT_TILES + j;
        bool s = false;
        uint32_t pix = 0, v = 0;
        if (tile < a.n_tiles) {
            v = collect_view(a, tile);
            const DvrView &V = a.views ? a.views[v] : a.one;
            uint8_t *const sf = V.stop;
            const uint32_t t = tile - V.stop_tile_off, tx = (V.width + 7u) / 8u;
            const uint32_t x = (t % tx) * 8u + (lane & 7u), y = (t / tx) * 8u + (lane >> 3);
            if (sf && x < V.width && y < V.height) {
                pix = y * V.width + x;
                s = sf[pix] != 0;
                if (s) sf[pix] = 0;
            }
        }
        masks[k] = __ballot(s);
        pixs[k] = pix;
        views[k] = v;
        if (lane == 0) {
            cnt[j] = (uint32_t)__popcll(masks[k]);
            vof[j] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t j = 0; j < COLLECT_TILES; ++j) total += cnt[j];
        base_s = total ? atomicAdd(a.out_count, total) : 0u;
        if (a.view_stopped && total) {
            // one atomic per run of tiles of the same view
            uint32_t run = 0, rv = vof[0];
            for (uint32_t j = 0; j < COLLECT_TILES; ++j) {
                if (vof[j] != rv) {
                    if (run) atomicAdd(&a.view_stopped[rv], run);
                    run = 0;
                    rv = vof[j];
                }
                run += cnt[j];
            }
            if (run) atomicAdd(&a.view_stopped[rv], run);
        }
    }
    __syncthreads();
    uint32_t off = base_s;
    for (uint32_t j = 0; j < wave * COLLECT_TILES_PER_WAVE; ++j) off += cnt[j];
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t k = 0; k < COLLECT_TILES_PER_WAVE; ++k) {
        if ((masks[k] >> lane) & 1ull)
            a.out_list[off + (uint32_t)__popcll(masks[k] & lt)] = (views[k] << DVR_PIXEL_BITS) | pixs[k];
        off += (uint32_t)__popcll(masks[k]);
    }
}

hipError_t launch_collect_stopped(const CollectArgs &a, hipStream_t stream)
{
    if (a.n_tiles == 0) return hipSuccess;
    k_collect_stopped<<<(a.n_tiles + COLLECT_TILES - 1u) / COLLECT_TILES, 256, 0, stream>>>(a);
    return hipGetLastError();
}

} // namespace gc
