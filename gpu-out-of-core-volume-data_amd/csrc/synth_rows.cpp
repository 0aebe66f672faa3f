// synth_rows.cpp -- the voxel loop of Synth::fill_box (synth.cpp), host only.
//
// Per row: the blob sum in blob order (one fused multiply-add per blob:
// sum = fma(w, fx, sum)), the noise, then rounding and clamping.  Every
// operation is element-wise IEEE single precision, each fma rounded once
// (built with -ffp-contract=off, so no other contraction), or 32/64-bit
// integer arithmetic, so the builds of this file for AVX-512, AVX2 and
// baseline x86-64 (SYNTH_ROWS_NAME names each one; the baseline one calls
// the C library's fmaf) write the same bytes.  The baseline build
// (SYNTH_ROWS_DISPATCH) also holds synth_rows(), which picks the widest build
// the host CPU runs (GC_SYNTH_ISA=base|avx2|avx512 in the environment picks
// one by name, for tests).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#if SYNTH_VEC_BYTES >= 32
#include <immintrin.h>
#endif

#include "synth.hpp"

#ifndef SYNTH_ROWS_NAME
#error "SYNTH_ROWS_NAME must name the build (synth_rows_avx512, synth_rows_avx2 or synth_rows_base)"
#endif

namespace gc {

// 32-bit avalanche hash (murmur3 finaliser), as Synth::noise_at uses it.
static inline uint32_t hash32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// value = clamp(round_half_up(sum) + noise) for one row, written as T
template <typename T>
static inline void store_row(const float *__restrict row, const int32_t *__restrict nz, int64_t x0, int64_t x1,
                             float fmax, T *__restrict out)
{
    for (int64_t kx = x0; kx < x1; ++kx) {
        float v = (float)(int32_t)(row[kx] + 0.5f) + (float)nz[kx];   // row >= 0: round half up
        v = v < 0.f ? 0.f : v;
        v = v > fmax ? fmax : v;
        if constexpr (sizeof(T) < 4)
            out[kx] = (T)(int32_t)v;   // v is a whole number in [0, fmax]: the same value
        else
            out[kx] = (T)v;
    }
}

// Vector width of this build (bytes) and the number of rows summed at once,
// so that the 32 x ROWS accumulators fill eight vector registers.
#ifndef SYNTH_VEC_BYTES
#define SYNTH_VEC_BYTES 16
#endif
constexpr int VL = SYNTH_VEC_BYTES / 4;   // floats per vector
constexpr int NV = 32 / VL;               // vectors per 32 voxels
constexpr int ROWS = 8 / NV > 0 ? 8 / NV : 1;
typedef float vf __attribute__((vector_size(SYNTH_VEC_BYTES)));

static inline vf vload(const float *p)
{
    vf v;
    std::memcpy(&v, p, sizeof(v));
    return v;
}

// c + w * f per element, rounded once
static inline vf vfma(float w, vf f, vf c)
{
#if SYNTH_VEC_BYTES == 64
    return (vf)_mm512_fmadd_ps(_mm512_set1_ps(w), (__m512)f, (__m512)c);
#elif SYNTH_VEC_BYTES == 32
    return (vf)_mm256_fmadd_ps(_mm256_set1_ps(w), (__m256)f, (__m256)c);
#else
    vf r;
    for (int i = 0; i < VL; ++i) r[i] = std::fma(w, f[i], c[i]);
    return r;
#endif
}

// Blob sums of R consecutive rows (ky0 ..), voxels [xb, xb+32) of each, into
// row[r * nx + xb ..]: 32 x R independent accumulators held in registers,
// each summing its blobs in blob order.  Weights and factors are >= 0, so a
// blob whose weight is 0 for a row gives fma(0, f, s) = s + 0 = s (s is never
// -0): no blob needs to be skipped.
template <int R>
static inline void sum_rows32(const SynthRows &a, uint32_t ky0, uint32_t xb, float *__restrict row)
{
    static_assert(R * NV <= 8, "eight accumulators");
    const uint32_t nx = a.n[0], ny = a.n[1];
    // accumulator j holds row j / NV, vector j % NV (named variables, so the
    // compiler keeps them in registers)
    vf c0{}, c1{}, c2{}, c3{}, c4{}, c5{}, c6{}, c7{};
    for (size_t k = 0; k < a.nb; ++k) {
        const float *w = &a.wz[k * ny + ky0];
        const float *f = &a.fx[k * nx + xb];
#define GC_ACC(j)                                                          \
        if constexpr ((j) < R * NV) c##j = vfma(w[(j) / NV], vload(f + ((j) % NV) * VL), c##j);
        GC_ACC(0) GC_ACC(1) GC_ACC(2) GC_ACC(3) GC_ACC(4) GC_ACC(5) GC_ACC(6) GC_ACC(7)
#undef GC_ACC
    }
#define GC_OUT(j)                                                                                    \
    if constexpr ((j) < R * NV) std::memcpy(row + ((j) / NV) * nx + xb + ((j) % NV) * VL, &c##j, sizeof(vf));
    GC_OUT(0) GC_OUT(1) GC_OUT(2) GC_OUT(3) GC_OUT(4) GC_OUT(5) GC_OUT(6) GC_OUT(7)
#undef GC_OUT
}

// NXC: the row width when it is known at compile time (32, the usual brick
// width), else 0; fixed trip counts let the short per-row loops run without
// run-time alias checks and remainders.
template <uint32_t NXC>
static void rows_impl(const SynthRows &a)
{
    const uint32_t nx = NXC ? NXC : a.n[0], ny = a.n[1], nzz = a.n[2];
    const uint32_t span = 2u * a.noise + 1u;
    const int32_t noise = (int32_t)a.noise;
    float *__restrict row = a.row;
    int32_t *__restrict nz = a.nz;
    const uint32_t *__restrict xkey = a.xkey;
    // x range inside the level
    const int64_t x0 = std::max<int64_t>(0, -a.o[0]),
                  x1 = std::min<int64_t>((int64_t)nx, (int64_t)a.ext[0] - a.o[0]);
    if (x1 <= x0) return;   // every voxel stays 0
    for (uint32_t kz = 0; kz < nzz; ++kz) {
        const int64_t gz = a.o[2] + kz;
        if (gz < 0 || gz >= (int64_t)a.ext[2]) continue;
        // the blob weights of this plane's rows: wz[k][ky] = amp fz(kz) fy(ky)
        for (size_t k = 0; k < a.nb; ++k) {
            const float fzk = a.fz[k * nzz + kz];
            for (uint32_t ky = 0; ky < ny; ++ky) a.wz[k * ny + ky] = fzk * a.fy[k * ny + ky];
        }
        for (uint32_t ky0 = 0; ky0 < ny; ky0 += ROWS) {
            const uint32_t nr = std::min((uint32_t)ROWS, ny - ky0);
            uint32_t xb = 0;
            if (nr == ROWS)
                for (; xb + 32 <= nx; xb += 32) sum_rows32<ROWS>(a, ky0, xb, row);
            else
                for (; xb + 32 <= nx; xb += 32)
                    for (uint32_t r = 0; r < nr; ++r) sum_rows32<1>(a, ky0 + r, xb, row + r * nx);
            if (xb < nx) {
                for (uint32_t r = 0; r < nr; ++r) {
                    float *rr = row + r * nx;
                    for (uint32_t i = xb; i < nx; ++i) rr[i] = 0.f;
                    for (size_t k = 0; k < a.nb; ++k) {
                        const float w = a.wz[k * ny + ky0 + r];
                        if (w == 0.f) continue;
                        const float *__restrict f = &a.fx[k * nx];
                        for (uint32_t i = xb; i < nx; ++i) rr[i] = std::fma(w, f[i], rr[i]);
                    }
                }
            }
            for (uint32_t r = 0; r < nr; ++r) {
                const uint32_t ky = ky0 + r;
                const int64_t gy = a.o[1] + ky;
                if (gy < 0 || gy >= (int64_t)a.ext[1]) continue;   // stays 0
                const uint64_t rbase = (uint64_t)nx * (ky + (uint64_t)ny * kz);
                if (noise) {
                    const uint32_t rk = a.seed ^ (a.level * 0x27d4eb2fu) ^ ((uint32_t)gz * 0x165667b1u) ^
                                        ((uint32_t)gy * 0xc2b2ae3du);
                    if (span <= 0xffffu) {
                        // (h * span) >> 32 in 32-bit lanes: with h = hi * 2^16 + lo,
                        // it equals (hi * span + ((lo * span) >> 16)) >> 16 exactly
                        for (uint32_t i = 0; i < nx; ++i) {
                            const uint32_t h = hash32(rk ^ xkey[i]);
                            const uint32_t q = ((h >> 16) * span + (((h & 0xffffu) * span) >> 16)) >> 16;
                            nz[i] = (int32_t)q - noise;
                        }
                    } else {
                        for (uint32_t i = 0; i < nx; ++i) {
                            const uint32_t h = hash32(rk ^ xkey[i]);
                            nz[i] = (int32_t)(uint32_t)(((uint64_t)h * span) >> 32) - noise;
                        }
                    }
                } else {
                    for (uint32_t i = 0; i < nx; ++i) nz[i] = 0;
                }
                const float *rr = row + r * nx;
                if (NXC && x0 == 0 && x1 == (int64_t)NXC) {
                    if (a.bytes == 1) store_row(rr, nz, 0, NXC, a.fmax, static_cast<uint8_t *>(a.dst) + rbase);
                    else if (a.bytes == 2) store_row(rr, nz, 0, NXC, a.fmax, static_cast<uint16_t *>(a.dst) + rbase);
                    else store_row(rr, nz, 0, NXC, a.fmax, static_cast<float *>(a.dst) + rbase);
                } else {
                    if (a.bytes == 1) store_row(rr, nz, x0, x1, a.fmax, static_cast<uint8_t *>(a.dst) + rbase);
                    else if (a.bytes == 2) store_row(rr, nz, x0, x1, a.fmax, static_cast<uint16_t *>(a.dst) + rbase);
                    else store_row(rr, nz, x0, x1, a.fmax, static_cast<float *>(a.dst) + rbase);
                }
            }
        }
    }
}

void SYNTH_ROWS_NAME(const SynthRows &a)
{
    if (a.n[0] == 32) rows_impl<32>(a);
    else rows_impl<0>(a);
}

#if defined(SYNTH_ROWS_DISPATCH)
void synth_rows(const SynthRows &a)
{
    static const auto fn = []() -> void (*)(const SynthRows &) {
        __builtin_cpu_init();
        if (const char *isa = std::getenv("GC_SYNTH_ISA")) {
            if (!std::strcmp(isa, "base")) return synth_rows_base;
            if (!std::strcmp(isa, "avx2") && __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma"))
                return synth_rows_avx2;
            if (!std::strcmp(isa, "avx512") && __builtin_cpu_supports("avx512f")) return synth_rows_avx512;
        }
        if (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
            __builtin_cpu_supports("avx512dq") && __builtin_cpu_supports("avx512vl"))
            return synth_rows_avx512;
        if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) return synth_rows_avx2;
        return synth_rows_base;
    }();
    fn(a);
}
#endif

} // namespace gc
