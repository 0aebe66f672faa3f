// lz4_block.cpp -- the LZ4 block format (the codec of the reference's brick
// files: BricksManager.cpp:93-114 compresses each big brick with
// LZ4_compress_HC into one block, :150-158 decodes it with
// LZ4_decompress_fast), written from the published block-format
// description:
//
//   block    = sequence* last
//   sequence = token [literal-length bytes] literals offset [match-length bytes]
//   last     = token [literal-length bytes] literals          (no match)
//   token    = literal length (high nibble) | match length - 4 (low nibble);
//              a nibble of 15 continues in bytes of 255 until one below 255
//   offset   = 2 bytes little endian, 1..65535 back from the current output
//
// with the end-of-block rules: the last 5 bytes are literals and the last
// match starts at least 12 bytes before the end, so blocks under 13 bytes
// are all literals.  The compressor is a greedy single-probe hash matcher:
// any valid block decodes with any LZ4 decoder, so its output need not be
// byte-identical to LZ4-HC's.
#include "gc_internal.hpp"

#include <vector>

namespace gc {

namespace {

constexpr uint64_t LZ4_MAX_INPUT = 0x7E000000ull;   // LZ4_MAX_INPUT_SIZE
constexpr uint32_t MIN_MATCH = 4;
constexpr uint64_t LAST_LITERALS = 5;
constexpr uint64_t MF_LIMIT = 12;
constexpr uint32_t HASH_BITS = 16;
constexpr uint32_t NO_POS = 0xffffffffu;

inline uint32_t read32(const uint8_t *p)
{
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}
inline uint32_t hash4(uint32_t v) { return (v * 2654435761u) >> (32 - HASH_BITS); }

// a length beyond a nibble's 15, in bytes of 255
inline bool put_length(uint8_t *dst, uint64_t cap, uint64_t &op, uint64_t extra)
{
    const uint64_t nb = extra / 255 + 1;
    if (op + nb > cap) return false;
    while (extra >= 255) {
        dst[op++] = 255;
        extra -= 255;
    }
    dst[op++] = (uint8_t)extra;
    return true;
}

inline bool put_literals(const uint8_t *src, uint64_t from, uint64_t n, uint8_t *dst, uint64_t cap, uint64_t &op,
                         uint32_t match_nibble)
{
    if (op + 1 > cap) return false;
    dst[op++] = (uint8_t)((n >= 15 ? 15u : (uint32_t)n) << 4 | match_nibble);
    if (n >= 15 && !put_length(dst, cap, op, n - 15)) return false;
    if (op + n > cap) return false;
    std::memcpy(dst + op, src + from, n);
    op += n;
    return true;
}

} // namespace

uint64_t lz4_bound(uint64_t n) { return n + n / 255 + 16; }

// Compressed size, or -1 when `cap` is too small or n too large.
int64_t lz4_compress(const uint8_t *src, uint64_t n, uint8_t *dst, uint64_t cap)
{
    if (n > LZ4_MAX_INPUT) return -1;
    uint64_t op = 0, anchor = 0;
    if (n >= MF_LIMIT + 1) {
        std::vector<uint32_t> table(1u << HASH_BITS, NO_POS);
        const uint64_t last_start = n - MF_LIMIT;      // a match starts at or before this
        const uint64_t match_end = n - LAST_LITERALS;  // and ends at or before this
        uint64_t ip = 0;
        while (ip <= last_start) {
            const uint32_t seq = read32(src + ip);
            const uint32_t h = hash4(seq);
            const uint32_t ref = table[h];
            table[h] = (uint32_t)ip;
            if (ref == NO_POS || ip - ref > 65535 || read32(src + ref) != seq) {
                ++ip;
                continue;
            }
            uint64_t m = ip, r = ref;
            // extend backwards into the pending literals
            while (m > anchor && r > 0 && src[m - 1] == src[r - 1]) {
                --m;
                --r;
            }
            uint64_t len = (ip - m) + MIN_MATCH;
            while (m + len < match_end && src[r + len] == src[m + len]) ++len;
            const uint64_t ml = len - MIN_MATCH;
            if (!put_literals(src, anchor, m - anchor, dst, cap, op, ml >= 15 ? 15u : (uint32_t)ml)) return -1;
            if (op + 2 > cap) return -1;
            const uint32_t off = (uint32_t)(m - r);
            dst[op++] = (uint8_t)(off & 0xff);
            dst[op++] = (uint8_t)(off >> 8);
            if (ml >= 15 && !put_length(dst, cap, op, ml - 15)) return -1;
            ip = m + len;
            anchor = ip;
            // positions inside the match seed later matches
            if (ip - 2 <= last_start) table[hash4(read32(src + ip - 2))] = (uint32_t)(ip - 2);
        }
    }
    if (!put_literals(src, anchor, n - anchor, dst, cap, op, 0)) return -1;
    return (int64_t)op;
}

// Decoded size (== dst_n), or -1 for a block that is malformed or does not
// decode to exactly dst_n bytes.
int64_t lz4_decompress(const uint8_t *src, uint64_t n, uint8_t *dst, uint64_t dst_n)
{
    uint64_t ip = 0, op = 0;
    for (;;) {
        if (ip >= n) return -1;
        const uint32_t token = src[ip++];
        uint64_t lit = token >> 4;
        if (lit == 15) {
            uint32_t b;
            do {
                if (ip >= n) return -1;
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (lit > n - ip || lit > dst_n - op) return -1;
        std::memcpy(dst + op, src + ip, lit);
        ip += lit;
        op += lit;
        if (ip == n) break;   // the last sequence has no match
        if (n - ip < 2) return -1;
        const uint64_t off = (uint64_t)src[ip] | (uint64_t)src[ip + 1] << 8;
        ip += 2;
        if (off == 0 || off > op) return -1;
        uint64_t ml = token & 15u;
        if (ml == 15) {
            uint32_t b;
            do {
                if (ip >= n) return -1;
                b = src[ip++];
                ml += b;
            } while (b == 255);
        }
        ml += MIN_MATCH;
        if (ml > dst_n - op) return -1;
        uint8_t *d = dst + op;
        const uint8_t *s = d - off;
        uint64_t k = 0;
        if (off >= 8)
            for (; k + 8 <= ml; k += 8) std::memcpy(d + k, s + k, 8);   // chunks never overlap their source
        for (; k < ml; ++k) d[k] = s[k];                                // overlapping (repeating) tail
        op += ml;
    }
    return op == dst_n ? (int64_t)op : -1;
}

} // namespace gc

using namespace gc;

extern "C" uint64_t gc_lz4_bound(uint64_t n) { return lz4_bound(n); }

extern "C" int gc_lz4_compress(const void *src, uint64_t n, void *dst, uint64_t cap, uint64_t *out_n)
{
    if ((!src && n) || !dst || !out_n) return fail(GC_ERR_ARG, "null argument");
    const int64_t r = lz4_compress(static_cast<const uint8_t *>(src), n, static_cast<uint8_t *>(dst), cap);
    if (r < 0) return fail(GC_ERR_ARG, "gc_lz4_compress: output capacity below gc_lz4_bound or input above 2,113,929,216 bytes");
    *out_n = (uint64_t)r;
    return GC_OK;
}

extern "C" int gc_lz4_decompress(const void *src, uint64_t n, void *dst, uint64_t dst_n)
{
    if (!src || (!dst && dst_n)) return fail(GC_ERR_ARG, "null argument");
    if (lz4_decompress(static_cast<const uint8_t *>(src), n, static_cast<uint8_t *>(dst), dst_n) < 0)
        return fail(GC_ERR_SOURCE, "gc_lz4_decompress: malformed LZ4 block or wrong decoded size");
    return GC_OK;
}
