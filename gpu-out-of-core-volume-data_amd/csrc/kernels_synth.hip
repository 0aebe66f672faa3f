// kernels_synth.hip -- seeded synthetic bricks written on the device.
//
// The device form of Synth::fill_box (synth.cpp / synth_rows.cpp): one
// workgroup per requested brick writes the brick straight into device memory
// (the cache's staging buffer), so a volume synthesised on demand costs no
// host arithmetic and no PCIe upload.  Per voxel inside the level:
//   sum  = sum over blobs k, in blob order, of (amp_k fz_k(z) * fy_k(y)) * fx_k(x),
//          one fused multiply-add per blob: sum = fma(w_k, fx_k(x), sum)
//   value = clamp((float)(int)(sum + 0.5f) + noise(x, y, z), 0, max)
// with the separable factors read from the host's per-level axis tables
// (copied to the device once), each product and fma a single IEEE float
// operation (the library is built with -ffp-contract=off) and the same
// integer noise hash, so every brick is bit-identical to the host's.  Blobs
// that do not reach the brick have a zero factor on some axis and add +0,
// which leaves the (non-negative) sum unchanged: the host skips them, the
// device adds them.  Voxels outside the level are 0.
#include <hip/hip_runtime.h>

#include "gc_internal.hpp"

namespace gc {

__device__ __forceinline__ uint32_t synth_hash32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

template <typename T>
__device__ __forceinline__ void synth_store(uint8_t *out, uint64_t i, float v)
{
    if constexpr (sizeof(T) < 4)
        reinterpret_cast<T *>(out)[i] = (T)(int32_t)v;
    else
        reinterpret_cast<float *>(out)[i] = v;
}

// Grid: one workgroup of 256 lanes per brick (p.n workgroups).  Dynamic LDS:
// 3 * nb * bs floats (the brick's x, y and amp-scaled z factors of every
// blob).
template <typename T>
__global__ __launch_bounds__(256) void k_synth_bricks(SynthLaunch p)
{
    extern __shared__ float sf[];
    const uint32_t u = blockIdx.x;
    const uint32_t *key = p.keys + 4ull * u;
    const uint32_t level = key[3], bs = p.bs, nb = p.nb, inner = bs - 2 * p.cov;
    if (level > p.max_level) return;   // the whole workgroup: another writer's brick
    const int64_t o0 = (int64_t)key[0] * inner - p.cov, o1 = (int64_t)key[1] * inner - p.cov,
                  o2 = (int64_t)key[2] * inner - p.cov;
    const uint32_t nf = nb * bs;
    for (uint32_t i = threadIdx.x; i < 3 * nf; i += blockDim.x) {
        const uint32_t a = i / nf, r = i - a * nf, k = r / bs, j = r - k * bs;
        const SynthAxis ax = p.axis[level][a];
        const int64_t c = (a == 0 ? o0 : (a == 1 ? o1 : o2)) + j;
        float f = 0.f;
        if (c >= ax.lo && c < ax.lo + (int64_t)ax.len) f = p.tab[ax.off + (uint64_t)k * ax.len + (uint64_t)(c - ax.lo)];
        if (a == 2) f = p.amp[k] * f;
        sf[i] = f;
    }
    __syncthreads();
    const float *fx = sf, *fy = sf + nf, *fz = sf + 2 * nf;
    const uint32_t e0 = p.ext[level][0], e1 = p.ext[level][1], e2 = p.ext[level][2];
    const uint32_t span = 2u * p.noise + 1u;
    const uint64_t nvox = (uint64_t)bs * bs * bs;
    uint8_t *out = p.dst + (uint64_t)u * nvox * sizeof(T);
    int above = 0;
    for (uint32_t v = threadIdx.x; v < nvox; v += blockDim.x) {
        const uint32_t x = v % bs, y = (v / bs) % bs, z = v / (bs * bs);
        const int64_t gx = o0 + x, gy = o1 + y, gz = o2 + z;
        float val = 0.f;
        if (gx >= 0 && gx < (int64_t)e0 && gy >= 0 && gy < (int64_t)e1 && gz >= 0 && gz < (int64_t)e2) {
            float c = 0.f;
            for (uint32_t k = 0; k < nb; ++k) {
                const float w = fz[k * bs + z] * fy[k * bs + y];
                c = __builtin_fmaf(w, fx[k * bs + x], c);
            }
            int32_t nz = 0;
            if (p.noise) {
                const uint32_t rk = p.seed ^ (level * 0x27d4eb2fu) ^ ((uint32_t)gz * 0x165667b1u) ^
                                    ((uint32_t)gy * 0xc2b2ae3du);
                const uint32_t h = synth_hash32(rk ^ ((uint32_t)gx * 0x9e3779b1u));
                nz = (int32_t)(uint32_t)(((uint64_t)h * span) >> 32) - (int32_t)p.noise;
            }
            float r = (float)(int32_t)(c + 0.5f) + (float)nz;
            r = r < 0.f ? 0.f : r;
            r = r > p.fmax ? p.fmax : r;
            val = r;
        }
        synth_store<T>(out, v, val);
        above |= !(val < p.thr);
    }
    // empty classification: every voxel below the threshold
    // (BrickProcessorPredicate.hpp:11-45)
    above = __syncthreads_or(above);
    if (threadIdx.x == 0) p.below[u] = above ? 0u : 1u;
}

// Bricks whose size is a multiple of 4: each lane writes four consecutive
// voxels of a row.  Per blob the row weight w_k = (amp_k fz_k(z)) * fy_k(y) is
// formed once for the four voxels and the four x factors come in one 16-byte
// LDS read, so a blob costs a lane three LDS reads per four voxels instead of
// three per voxel (the per-voxel kernel above is bound by LDS issue).  Every
// voxel still sums fma(w_k, fx_k(x), sum) in blob order from the same
// products, so the bytes are those of k_synth_bricks.
template <typename T>
__global__ __launch_bounds__(256) void k_synth_bricks_x4(SynthLaunch p)
{
    extern __shared__ __align__(16) float sf4[];
    const uint32_t u = blockIdx.x;
    const uint32_t *key = p.keys + 4ull * u;
    const uint32_t level = key[3], bs = p.bs, nb = p.nb, inner = bs - 2 * p.cov;
    if (level > p.max_level) return;   // the whole workgroup: another writer's brick
    const int64_t o0 = (int64_t)key[0] * inner - p.cov, o1 = (int64_t)key[1] * inner - p.cov,
                  o2 = (int64_t)key[2] * inner - p.cov;
    const uint32_t nf = nb * bs;
    for (uint32_t i = threadIdx.x; i < 3 * nf; i += blockDim.x) {
        const uint32_t a = i / nf, r = i - a * nf, k = r / bs, j = r - k * bs;
        const SynthAxis ax = p.axis[level][a];
        const int64_t c = (a == 0 ? o0 : (a == 1 ? o1 : o2)) + j;
        float f = 0.f;
        if (c >= ax.lo && c < ax.lo + (int64_t)ax.len) f = p.tab[ax.off + (uint64_t)k * ax.len + (uint64_t)(c - ax.lo)];
        if (a == 2) f = p.amp[k] * f;
        sf4[i] = f;
    }
    __syncthreads();
    const float *fx = sf4, *fy = sf4 + nf, *fz = sf4 + 2 * nf;
    const int64_t e0 = p.ext[level][0], e1 = p.ext[level][1], e2 = p.ext[level][2];
    const uint32_t span = 2u * p.noise + 1u;
    const uint32_t qx = bs / 4u, nq = qx * bs * bs;
    T *out = reinterpret_cast<T *>(p.dst + (uint64_t)u * bs * bs * bs * sizeof(T));
    int above = 0;
    for (uint32_t q = threadIdx.x; q < nq; q += blockDim.x) {
        const uint32_t x = (q % qx) * 4u, r = q / qx, y = r % bs, z = r / bs;
        const int64_t gx0 = o0 + x, gy = o1 + y, gz = o2 + z;
        const bool yz_in = gy >= 0 && gy < e1 && gz >= 0 && gz < e2;
        float c0 = 0.f, c1 = 0.f, c2 = 0.f, c3 = 0.f;
        if (yz_in) {
            for (uint32_t k = 0; k < nb; ++k) {
                const float w = fz[k * bs + z] * fy[k * bs + y];
                const float4 f = *reinterpret_cast<const float4 *>(fx + k * bs + x);
                c0 = __builtin_fmaf(w, f.x, c0);
                c1 = __builtin_fmaf(w, f.y, c1);
                c2 = __builtin_fmaf(w, f.z, c2);
                c3 = __builtin_fmaf(w, f.w, c3);
            }
        }
        const uint32_t rk = p.seed ^ (level * 0x27d4eb2fu) ^ ((uint32_t)gz * 0x165667b1u) ^ ((uint32_t)gy * 0xc2b2ae3du);
        const float cs[4] = {c0, c1, c2, c3};
        T vals[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t gx = gx0 + i;
            float val = 0.f;
            if (yz_in && gx >= 0 && gx < e0) {
                int32_t nz = 0;
                if (p.noise) {
                    const uint32_t h = synth_hash32(rk ^ ((uint32_t)gx * 0x9e3779b1u));
                    nz = (int32_t)(uint32_t)(((uint64_t)h * span) >> 32) - (int32_t)p.noise;
                }
                float rr = (float)(int32_t)(cs[i] + 0.5f) + (float)nz;
                rr = rr < 0.f ? 0.f : rr;
                rr = rr > p.fmax ? p.fmax : rr;
                val = rr;
            }
            above |= !(val < p.thr);
            if constexpr (sizeof(T) < 4) vals[i] = (T)(int32_t)val;
            else vals[i] = val;
        }
        // four consecutive voxels, one aligned store (x and bs are multiples of 4)
        T *o = out + x + (uint64_t)bs * (y + (uint64_t)bs * z);
        if constexpr (sizeof(T) == 1) {
            *reinterpret_cast<uint32_t *>(o) = (uint32_t)vals[0] | ((uint32_t)vals[1] << 8) |
                                               ((uint32_t)vals[2] << 16) | ((uint32_t)vals[3] << 24);
        } else if constexpr (sizeof(T) == 2) {
            *reinterpret_cast<uint2 *>(o) = make_uint2((uint32_t)vals[0] | ((uint32_t)vals[1] << 16),
                                                       (uint32_t)vals[2] | ((uint32_t)vals[3] << 16));
        } else {
            *reinterpret_cast<float4 *>(o) = make_float4(vals[0], vals[1], vals[2], vals[3]);
        }
    }
    above = __syncthreads_or(above);
    if (threadIdx.x == 0) p.below[u] = above ? 0u : 1u;
}

// A box of voxels (x fastest) of one level of the synthetic volume, the
// values the bricks of k_synth_bricks hold at the same positions.  A
// workgroup writes a 256-voxel segment of four rows (y .. y + 3): the rows
// share the segment's x factors, so a lane reads each factor once for its
// four voxels (beside the four row weights w_k = (amp_k fz_k(z)) * fy_k(y),
// the same for every lane), and every voxel sums fma(w_k, fx_k(x), sum) in
// blob order as the brick kernel does.  Grid: (ceil(n[0] / 256),
// ceil(n[1] / 4), n[2]).  Dynamic LDS: nb x 256 x factors, then 4 x nb row
// weights.
template <typename T>
__global__ __launch_bounds__(256) void k_synth_box_r4(SynthBox b)
{
    extern __shared__ float fxr[];
    const uint32_t nb = b.nb;
    float *ws = fxr + nb * 256u;   // ws[r * nb + k]
    const uint32_t x0 = blockIdx.x * 256u, y0 = blockIdx.y * 4u, z = blockIdx.z;
    const int64_t gz = b.o[2] + z;
    for (uint32_t i = threadIdx.x; i < 4u * nb; i += blockDim.x) {
        const uint32_t r = i / nb, k = i - r * nb;
        const int64_t gy = b.o[1] + y0 + r;
        const SynthAxis &ay = b.axis[1], &az = b.axis[2];
        const float fy = gy >= ay.lo && gy < ay.lo + (int64_t)ay.len ? b.tab[ay.off + (uint64_t)k * ay.len + (uint64_t)(gy - ay.lo)] : 0.f;
        const float fz0 = gz >= az.lo && gz < az.lo + (int64_t)az.len ? b.tab[az.off + (uint64_t)k * az.len + (uint64_t)(gz - az.lo)] : 0.f;
        const float fz = b.amp[k] * fz0;
        ws[i] = fz * fy;
    }
    for (uint32_t i = threadIdx.x; i < nb * 256u; i += blockDim.x) {
        const uint32_t k = i / 256u, j = i % 256u;
        const int64_t c = b.o[0] + x0 + j;
        const SynthAxis &ax = b.axis[0];
        float f = 0.f;
        if (c >= ax.lo && c < ax.lo + (int64_t)ax.len) f = b.tab[ax.off + (uint64_t)k * ax.len + (uint64_t)(c - ax.lo)];
        fxr[i] = f;
    }
    __syncthreads();
    const uint32_t x = x0 + threadIdx.x;
    if (x >= b.n[0]) return;
    const int64_t gx = b.o[0] + x;
    float c0 = 0.f, c1 = 0.f, c2 = 0.f, c3 = 0.f;
    for (uint32_t k = 0; k < nb; ++k) {
        const float f = fxr[k * 256u + threadIdx.x];
        c0 = __builtin_fmaf(ws[k], f, c0);
        c1 = __builtin_fmaf(ws[nb + k], f, c1);
        c2 = __builtin_fmaf(ws[2u * nb + k], f, c2);
        c3 = __builtin_fmaf(ws[3u * nb + k], f, c3);
    }
    const float cs[4] = {c0, c1, c2, c3};
#pragma unroll
    for (uint32_t r = 0; r < 4u; ++r) {
        const uint32_t y = y0 + r;
        if (y >= b.n[1]) break;
        const int64_t gy = b.o[1] + y;
        float val = 0.f;
        if (gx >= 0 && gx < (int64_t)b.ext[0] && gy >= 0 && gy < (int64_t)b.ext[1] && gz >= 0 && gz < (int64_t)b.ext[2]) {
            int32_t nz = 0;
            if (b.noise) {
                const uint32_t span = 2u * b.noise + 1u;
                const uint32_t rk = b.seed ^ (b.level * 0x27d4eb2fu) ^ ((uint32_t)gz * 0x165667b1u) ^ ((uint32_t)gy * 0xc2b2ae3du);
                const uint32_t h = synth_hash32(rk ^ ((uint32_t)gx * 0x9e3779b1u));
                nz = (int32_t)(uint32_t)(((uint64_t)h * span) >> 32) - (int32_t)b.noise;
            }
            float rr = (float)(int32_t)(cs[r] + 0.5f) + (float)nz;
            rr = rr < 0.f ? 0.f : rr;
            rr = rr > b.fmax ? b.fmax : rr;
            val = rr;
        }
        synth_store<T>(b.dst, (uint64_t)x + (uint64_t)b.n[0] * ((uint64_t)y + (uint64_t)b.n[1] * z), val);
    }
}

hipError_t launch_synth_box(const SynthBox &b, hipStream_t s)
{
    if (b.n[0] == 0 || b.n[1] == 0 || b.n[2] == 0) return hipSuccess;
    if (b.n[2] > 65535u || (b.n[1] + 3u) / 4u > 65535u) return hipErrorInvalidValue;
    const dim3 grid((b.n[0] + 255u) / 256u, (b.n[1] + 3u) / 4u, b.n[2]);
    const size_t lds = (size_t)b.nb * (256u + 4u) * sizeof(float);
    if (b.bytes == 1) k_synth_box_r4<uint8_t><<<grid, 256, lds, s>>>(b);
    else if (b.bytes == 2) k_synth_box_r4<uint16_t><<<grid, 256, lds, s>>>(b);
    else k_synth_box_r4<float><<<grid, 256, lds, s>>>(b);
    return hipGetLastError();
}

hipError_t launch_synth_bricks(const SynthLaunch &p, hipStream_t s)
{
    if (p.n == 0) return hipSuccess;
    const size_t lds = (size_t)3 * p.nb * p.bs * sizeof(float);
    if (p.bs % 4u == 0) {
        if (p.bytes == 1) k_synth_bricks_x4<uint8_t><<<p.n, 256, lds, s>>>(p);
        else if (p.bytes == 2) k_synth_bricks_x4<uint16_t><<<p.n, 256, lds, s>>>(p);
        else k_synth_bricks_x4<float><<<p.n, 256, lds, s>>>(p);
        return hipGetLastError();
    }
    if (p.bytes == 1) k_synth_bricks<uint8_t><<<p.n, 256, lds, s>>>(p);
    else if (p.bytes == 2) k_synth_bricks<uint16_t><<<p.n, 256, lds, s>>>(p);
    else k_synth_bricks<float><<<p.n, 256, lds, s>>>(p);
    return hipGetLastError();
}

} // namespace gc
