// synth.hpp -- host utilities of libgcvol.so (see synth.cpp).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "gcvol.h"

namespace gc {

uint64_t splitmix64(uint64_t &state);

// Seeded synthetic volume: Gaussian blobs plus uniform integer noise,
// rounded and clamped to the voxel type.  Level l is evaluated at the level-l
// voxel centres, so every level is synthesised directly (no mip needed).
struct Synth {
    struct Blob {
        double c[3];
        double k;        // 1 / (2 sigma^2)
        double amp;
        double radius;   // 4 sigma
    };
    gc_synth_params prm;
    uint32_t dims[3];
    uint32_t bytes;
    double maxval;
    std::vector<Blob> blobs;
    // Separable factors of every blob at every voxel coordinate of a level,
    // computed once: table[l][a] holds blob k's factor along axis a at level-l
    // coordinate i at [k * len + i - lo], for i in [lo, lo + len) (the level
    // extent plus TABLE_PAD voxels on each side).  Coordinates outside it are
    // evaluated directly; both give the same float.
    static constexpr int64_t TABLE_PAD = 64;
    struct AxisTable {
        int64_t lo = 0, len = 0;
        std::vector<float> f;
    };
    std::vector<std::array<AxisTable, 3>> table;

    Synth(const gc_synth_params &p, const uint32_t dims0[3], uint32_t bytes);
    int noise_at(uint32_t level, int64_t x, int64_t y, int64_t z) const;
    void build_tables();
    void axis_factors(size_t k, int a, uint32_t level, int64_t o, uint32_t n, float scale, float *out) const;
    void fill_box(uint32_t level, const int64_t o[3], const uint32_t n[3], const uint32_t ext[3],
                  void *dst) const;
};

// Arguments of the voxel loop of Synth::fill_box: separable blob factors
// fx[k][i], fy[k][j], amp*fz[k][l] of the nb blobs that reach the box, the
// x keys of the noise hash and scratch for 4 rows of sums (row, 4*n[0]
// floats) and one row of noise (nz, n[0] entries).
struct SynthRows {
    uint32_t level, seed, noise, bytes;
    float fmax;
    size_t nb;
    const int64_t *o;
    const uint32_t *n, *ext;
    const float *fx, *fy, *fz;
    const uint32_t *xkey;
    float *row;
    int32_t *nz;
    float *wz;   // scratch, nb * n[1] floats
    void *dst;
};
// synth_rows.cpp is compiled three times (AVX-512, AVX2, baseline x86-64);
// synth_rows() picks the widest one the host CPU runs.
void synth_rows(const SynthRows &a);
void synth_rows_avx512(const SynthRows &a);
void synth_rows_avx2(const SynthRows &a);
void synth_rows_base(const SynthRows &a);

void tf_samples(const float *knots, uint32_t nk, uint32_t n, float *out);
void tf_seeded(uint64_t seed, uint32_t nk, float *knots);
void camera_matrix(float rx, float ry, const float t[3], float out[12]);
uint32_t level_count(const uint32_t dims[3], uint32_t bs);
void level_sizes(const uint32_t dims[3], uint32_t nl, uint32_t (*out)[3]);
void real_sizes(uint32_t nl, const uint32_t (*levels)[3], uint32_t bs, uint32_t cov, const uint32_t bb[3],
                uint32_t (*out)[3]);
void pool_shape(uint64_t budget, uint64_t brick_bytes, uint32_t out[3]);
bool cfg_lookup(const std::string &path, const std::string &key, std::string &value);

} // namespace gc
