// synth.cpp -- host utilities of libgcvol.so: seeded synthetic volumes and
// bricks, transfer-function sampling, the reference viewer's camera, the
// .cfg reader and the preprocessing metadata rules.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "gc_internal.hpp"
#include "synth.hpp"

namespace gc {

// ------------------------------------------------------------ hashing / PRNG
uint64_t splitmix64(uint64_t &state)
{
    uint64_t z = (state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static inline double uniform01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

Synth::Synth(const gc_synth_params &p, const uint32_t dims0[3], uint32_t bytes) : prm(p), bytes(bytes)
{
    for (int a = 0; a < 3; ++a) dims[a] = dims0[a];
    maxval = bytes == 1 ? 255.0 : (bytes == 2 ? 65535.0 : 1.0e30);
    uint64_t s = p.seed;
    blobs.resize(p.nb_blobs);
    for (auto &b : blobs) {
        b.c[0] = uniform01(s) * dims[0];
        b.c[1] = uniform01(s) * dims[1];
        b.c[2] = uniform01(s) * dims[2];
        const double sigma = p.sigma_min + uniform01(s) * (p.sigma_max - p.sigma_min);
        b.amp = p.amp_min + uniform01(s) * (p.amp_max - p.amp_min);
        b.k = 1.0 / (2.0 * sigma * sigma);
        b.radius = 4.0 * sigma;   // contributions beyond 4 sigma are exactly 0
    }
    build_tables();
}

// Separable Gaussian factor along one axis for a level-l voxel coordinate:
// the voxel centre in level-0 units is (i + 0.5) * 2^l.  It depends on the
// global coordinate only, so a voxel gets the same value in every box.
static inline float axis_factor(const Synth::Blob &b, int a, uint32_t level, int64_t i)
{
    const double c = ((double)i + 0.5) * (double)(1u << level);
    const double d = c - b.c[a];
    if (std::fabs(d) > b.radius) return 0.f;
    return (float)std::exp(-d * d * b.k);
}

void Synth::build_tables()
{
    const size_t nb = blobs.size();
    for (uint32_t l = 0; l < 16; ++l) {
        bool any = false;
        std::array<AxisTable, 3> t;
        for (int a = 0; a < 3; ++a) {
            const int64_t ext = ((int64_t)dims[a] + ((int64_t)1 << l) - 1) >> l;
            any = any || ext > 1 || l == 0;
            t[a].lo = -TABLE_PAD;
            t[a].len = ext + 2 * TABLE_PAD;
            t[a].f.resize(nb * (size_t)t[a].len);
            for (size_t k = 0; k < nb; ++k)
                for (int64_t i = 0; i < t[a].len; ++i)
                    t[a].f[k * (size_t)t[a].len + (size_t)i] = axis_factor(blobs[k], a, l, t[a].lo + i);
        }
        table.push_back(std::move(t));
        if (!any) break;   // every axis is down to one voxel
    }
}

// Factors of blob k along axis a for level-l coordinates [o, o + n): from
// the level's table when the range lies inside it.
void Synth::axis_factors(size_t k, int a, uint32_t level, int64_t o, uint32_t n, float scale, float *out) const
{
    if (level < table.size()) {
        const AxisTable &t = table[level][a];
        if (o >= t.lo && o + (int64_t)n <= t.lo + t.len) {
            const float *src = &t.f[k * (size_t)t.len + (size_t)(o - t.lo)];
            if (scale == 1.f)
                for (uint32_t i = 0; i < n; ++i) out[i] = src[i];
            else
                for (uint32_t i = 0; i < n; ++i) out[i] = scale * src[i];
            return;
        }
    }
    for (uint32_t i = 0; i < n; ++i) {
        const float f = axis_factor(blobs[k], a, level, o + i);
        out[i] = scale == 1.f ? f : scale * f;
    }
}

// 32-bit avalanche hash (murmur3 finaliser) of a voxel key.
static inline uint32_t hash32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

int Synth::noise_at(uint32_t level, int64_t x, int64_t y, int64_t z) const
{
    if (prm.noise == 0) return 0;
    const uint32_t span = 2u * prm.noise + 1u;
    const uint32_t key = (uint32_t)prm.seed ^ ((uint32_t)level * 0x27d4eb2fu) ^ ((uint32_t)z * 0x165667b1u) ^
                         ((uint32_t)y * 0xc2b2ae3du);
    const uint32_t h = hash32(key ^ ((uint32_t)x * 0x9e3779b1u));
    return (int)(((uint64_t)h * span) >> 32) - (int)prm.noise;
}

// Fill a box [o, o+n) of level-`level` voxels (x-fastest into dst, row pitch
// n[0]); voxels outside the level extent `ext` are 0.  value = clamp(
// round(sum of blobs) + noise), with the blob sum in float in blob order, one
// fused multiply-add per blob.
void Synth::fill_box(uint32_t level, const int64_t o[3], const uint32_t n[3], const uint32_t ext[3],
                     void *dst) const
{
    // blobs that reach the box (the others contribute exact zeros)
    std::vector<const Blob *> near;
    const double sc = (double)(1u << level);
    for (const auto &b : blobs) {
        bool ok = true;
        for (int a = 0; a < 3 && ok; ++a) {
            const double lo = ((double)o[a] + 0.5) * sc, hi = ((double)(o[a] + n[a] - 1) + 0.5) * sc;
            if (b.c[a] + b.radius < lo || b.c[a] - b.radius > hi) ok = false;
        }
        if (ok) near.push_back(&b);
    }
    const size_t nb = near.size();
    std::vector<float> fx(nb * n[0]), fy(nb * n[1]), fz(nb * n[2]);
    for (size_t k = 0; k < nb; ++k) {
        const size_t bk = (size_t)(near[k] - blobs.data());
        axis_factors(bk, 0, level, o[0], n[0], 1.f, &fx[k * n[0]]);
        axis_factors(bk, 1, level, o[1], n[1], 1.f, &fy[k * n[1]]);
        axis_factors(bk, 2, level, o[2], n[2], (float)near[k]->amp, &fz[k * n[2]]);
    }
    // per-row x keys of the noise hash
    std::vector<uint32_t> xkey(n[0]);
    for (uint32_t i = 0; i < n[0]; ++i) xkey[i] = (uint32_t)(o[0] + i) * 0x9e3779b1u;
    SynthRows a;
    a.level = level;
    a.seed = (uint32_t)prm.seed;
    a.noise = prm.noise;
    a.bytes = bytes;
    a.fmax = (float)maxval;
    a.nb = nb;
    a.o = o;
    a.n = n;
    a.ext = ext;
    a.fx = fx.data();
    a.fy = fy.data();
    a.fz = fz.data();
    a.xkey = xkey.data();
    std::vector<float> row(4 * (size_t)n[0]);
    std::vector<int32_t> nz(n[0]);
    a.row = row.data();
    a.nz = nz.data();
    std::vector<float> wz(nb * n[1]);
    a.wz = wz.data();
    a.dst = dst;
    synth_rows(a);
}

// -------------------------------------------------------- transfer function
void tf_samples(const float *knots, uint32_t nk, uint32_t n, float *out)
{
    // TransferFunction::regenerateSamples / getSample (TransferFunction.cpp:121-157)
    for (uint32_t i = 0; i < n; ++i) {
        const float x = (float)i / (float)n;
        float *o = out + 4 * i;
        if (nk == 0) { o[0] = o[1] = o[2] = o[3] = 0.f; continue; }
        uint32_t pos = 0;
        while (pos < nk && !(knots[5 * pos] > x)) ++pos;
        float prev[5], next[5];
        if (pos == 0) { prev[0] = 0.f; prev[1] = 0.f; std::memcpy(prev + 2, knots + 2, 3 * sizeof(float)); }
        else std::memcpy(prev, knots + 5 * (pos - 1), 5 * sizeof(float));
        if (pos == nk) { next[0] = 1.f; next[1] = 0.f; std::memcpy(next + 2, knots + 5 * (nk - 1) + 2, 3 * sizeof(float)); }
        else std::memcpy(next, knots + 5 * pos, 5 * sizeof(float));
        const float t = (x - prev[0]) / (next[0] - prev[0]);
        o[0] = prev[2] + (next[2] - prev[2]) * t;
        o[1] = prev[3] + (next[3] - prev[3]) * t;
        o[2] = prev[4] + (next[4] - prev[4]) * t;
        o[3] = prev[1] + (next[1] - prev[1]) * t;
    }
}

void tf_seeded(uint64_t seed, uint32_t nk, float *knots)
{
    // nk control points: increasing values in (0.05, 0.95), alpha and colour
    // uniform in [0, 1]; the first knot starts transparent.
    uint64_t s = seed ^ 0x7f4a7c159e3779b9ull;
    std::vector<float> vals(nk);
    for (uint32_t i = 0; i < nk; ++i) vals[i] = (float)(0.05 + 0.9 * uniform01(s));
    std::sort(vals.begin(), vals.end());
    for (uint32_t i = 0; i < nk; ++i) {
        knots[5 * i + 0] = vals[i];
        knots[5 * i + 1] = i == 0 ? 0.f : (float)(0.05 + 0.95 * uniform01(s));
        knots[5 * i + 2] = (float)uniform01(s);
        knots[5 * i + 3] = (float)uniform01(s);
        knots[5 * i + 4] = (float)uniform01(s);
    }
}

// ------------------------------------------------------------------- camera
void camera_matrix(float rx, float ry, const float t[3], float out[12])
{
    // updateViewMatrix (Display_helper.cpp:93-118): M = Rx(-rx) * Ry(-ry) *
    // T(-t); out holds its first three rows.
    const double ax = -rx * M_PI / 180.0, ay = -ry * M_PI / 180.0;
    const double cx = std::cos(ax), sx = std::sin(ax), cy = std::cos(ay), sy = std::sin(ay);
    const double Rx[3][3] = {{1, 0, 0}, {0, cx, -sx}, {0, sx, cx}};
    const double Ry[3][3] = {{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}};
    double R[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            R[i][j] = 0;
            for (int k = 0; k < 3; ++k) R[i][j] += Rx[i][k] * Ry[k][j];
        }
    for (int i = 0; i < 3; ++i) {
        double tr = 0;
        for (int k = 0; k < 3; ++k) tr += R[i][k] * (-(double)t[k]);
        for (int j = 0; j < 3; ++j) out[4 * i + j] = (float)R[i][j];
        out[4 * i + 3] = (float)tr;
    }
}

// ------------------------------------------------------------- config rules
uint32_t level_count(const uint32_t dims[3], uint32_t bs)
{
    // Mipmapper.cu:441-443
    const uint32_t m = std::max(dims[0], std::max(dims[1], dims[2]));
    return (uint32_t)std::ceil(std::log((double)(m / bs)) / std::log(2.0)) + 1;
}

void level_sizes(const uint32_t dims[3], uint32_t nl, uint32_t (*out)[3])
{
    // Mipmapper.cu:583-597 (ratio 2)
    uint32_t s[3] = {dims[0], dims[1], dims[2]};
    for (uint32_t i = 0; i < nl; ++i)
        for (int a = 0; a < 3; ++a) {
            out[i][a] = s[a];
            if (s[a] >= 2) s[a] /= 2;
        }
}

void real_sizes(uint32_t nl, const uint32_t (*levels)[3], uint32_t bs, uint32_t cov, const uint32_t bb[3],
                uint32_t (*out)[3])
{
    // Bricker_v2.cpp:45-76
    const uint32_t inner = bs - 2 * cov;
    for (uint32_t i = 0; i < nl; ++i)
        for (int a = 0; a < 3; ++a) {
            uint32_t nb = (uint32_t)std::ceil((float)levels[i][a] / (float)inner);
            nb = (uint32_t)std::ceil((float)nb / (float)bb[a]) * bb[a];
            out[i][a] = nb * inner;
        }
}

void pool_shape(uint64_t budget, uint64_t brick_bytes, uint32_t out[3])
{
    // Application.cu:85-93
    const uint64_t n = budget / brick_bytes;
    const uint64_t s = (uint64_t)std::cbrt((double)n);
    out[0] = out[1] = out[2] = (uint32_t)s;
    int d = 0;
    while ((uint64_t)out[0] * out[1] * out[2] < n) {
        out[d] += 1;
        d = (d + 1) % 3;
    }
}

bool cfg_lookup(const std::string &path, const std::string &key, std::string &value)
{
    // <Fields><Field key="..." value="..."/></Fields> (ConfigurationPolicies.cpp:16-44)
    std::ifstream f(path);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    std::string s = ss.str();
    // drop <!-- ... --> comments
    for (size_t c0; (c0 = s.find("<!--")) != std::string::npos;) {
        const size_t c1 = s.find("-->", c0 + 4);
        s.erase(c0, c1 == std::string::npos ? std::string::npos : c1 + 3 - c0);
    }
    size_t pos = 0;
    auto attr = [&](size_t from, size_t to, const char *name, std::string &out) {
        const std::string pat = std::string(name) + "=\"";
        size_t a = s.find(pat, from);
        if (a == std::string::npos || a > to) return false;
        a += pat.size();
        const size_t b = s.find('"', a);
        if (b == std::string::npos || b > to) return false;
        out = s.substr(a, b - a);
        return true;
    };
    bool found = false;
    while ((pos = s.find("<Field", pos)) != std::string::npos) {
        const size_t end = s.find('>', pos);
        if (end == std::string::npos) break;
        std::string k, v;
        if (s.compare(pos, 7, "<Fields") != 0 && attr(pos, end, "key", k) && attr(pos, end, "value", v) && k == key) {
            value = v;   // the last occurrence wins (Configuration::add_field)
            found = true;
        }
        pos = end;
    }
    return found;
}

} // namespace gc
