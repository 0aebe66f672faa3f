"""The on-disk brick store (CPU): LZ4 block codec, Bkey packing, brick file
layout, bricks.xml and the store brick source with its host LRU
(BricksManager.cpp:29-227, MetaData.cpp:36-216, BrickKey.cpp:15-25,
Bricker_v2.cpp:79-252).

The LZ4 codec is checked against the system's LZ4 runtime (liblz4.so.1,
called through ctypes) in both directions when that library is present:
its decoder reads our blocks, and our decoder reads the blocks of its LZ4-HC
compressor at level 9, the call the reference's brick writer makes
(BricksManager.cpp:100-105).
"""
import ctypes as C
import os
import struct

import numpy as np
import pytest

import gcvol as G


def _system_lz4():
    try:
        L = C.CDLL("liblz4.so.1")
    except OSError:
        return None
    for name in ("LZ4_compress_HC", "LZ4_compress_default", "LZ4_decompress_safe", "LZ4_compressBound"):
        if not hasattr(L, name):
            return None
    L.LZ4_compressBound.argtypes = [C.c_int]
    L.LZ4_compress_HC.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int]
    L.LZ4_compress_default.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int]
    L.LZ4_decompress_safe.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int]
    return L


def _inputs():
    rng = np.random.default_rng(7)
    out = [b"", b"a", b"abc", bytes(range(12)), bytes(13), bytes(range(20)), b"ab" * 40,
           bytes(1 << 20), rng.integers(0, 256, 65536, dtype=np.uint8).tobytes(),
           rng.integers(0, 4, 100000, dtype=np.uint8).tobytes(),
           (b"0123456789abcdef" * 5000)[:77777]]
    # a smooth u16 field with noise, like a brick of the synthetic volumes
    z, y, x = np.mgrid[0:32, 0:32, 0:32]
    f = (3000 * np.sin(x / 5.0) * np.cos(y / 7.0) + 20 * z + 30000).astype(np.uint16)
    f += rng.integers(0, 64, f.shape).astype(np.uint16)
    out.append(f.tobytes())
    # long literal runs and long matches (length bytes of 255)
    out.append(rng.integers(0, 256, 5000, dtype=np.uint8).tobytes() + bytes(70000) +
               rng.integers(0, 256, 300, dtype=np.uint8).tobytes())
    return out


def test_lz4_round_trip(gcv):
    for data in _inputs():
        c = G.lz4_compress(data)
        assert len(c) <= G.lib().gc_lz4_bound(len(data))
        assert G.lz4_decompress(c, len(data)) == data, len(data)
    assert len(G.lz4_compress(bytes(1 << 20))) < 5000     # zeros compress


def test_lz4_interoperates_with_system_lz4(gcv):
    L = _system_lz4()
    if L is None:
        pytest.skip("no liblz4.so.1 on this machine")
    for data in _inputs():
        n = len(data)
        # the system decoder reads our blocks
        c = G.lz4_compress(data)
        out = C.create_string_buffer(max(n, 1))
        assert L.LZ4_decompress_safe(c, out, len(c), n) == n
        assert out.raw[:n] == data
        # our decoder reads LZ4-HC level 9 (the reference's writer) and default blocks
        cap = L.LZ4_compressBound(n)
        for hc in (True, False):
            buf = C.create_string_buffer(max(cap, 1))
            k = L.LZ4_compress_HC(data, buf, n, cap, 9) if hc else L.LZ4_compress_default(data, buf, n, cap)
            assert k > 0
            assert G.lz4_decompress(buf.raw[:k], n) == data, (n, hc)


def test_lz4_rejects_malformed_blocks(gcv):
    data = (b"0123456789" * 200)
    c = G.lz4_compress(data)
    L = G.lib()

    def dec(blk, n):
        src = np.frombuffer(blk, np.uint8) if blk else np.zeros(1, np.uint8)
        dst = np.zeros(max(n, 1), np.uint8)
        return L.gc_lz4_decompress(G.ptr(src), len(blk), G.ptr(dst), n)

    assert dec(c, len(data)) == 0
    assert dec(c, len(data) - 1) == 4 and dec(c, len(data) + 1) == 4      # wrong size
    assert dec(c[:-3], len(data)) == 4                                    # truncated
    assert dec(b"", 0) == 4                                               # no token
    assert dec(bytes([0x10, 65, 0, 0]), 8) == 4                           # offset 0
    assert dec(bytes([0x10, 65, 2, 0]), 8) == 4                           # offset beyond the output
    assert dec(bytes([0xf0]), 20) == 4                                    # length bytes missing


def test_brick_key_packing(gcv):
    rng = np.random.default_rng(3)
    for _ in range(200):
        l, x, y, z = (int(v) for v in rng.integers(0, 65536, 4))
        k = G.brick_key(l, x, y, z)
        assert k == struct.unpack("<Q", struct.pack("<4H", l, x, y, z))[0]
        assert G.brick_key_unpack(k) == (l, x, y, z)
    assert G.brick_key(3, 1, 2, 4) == 3 | 1 << 16 | 2 << 32 | 4 << 48
    assert G.store_brick_path("/data/vol/", 32, 2, (1, 8, 2)) == "/data/vol/bricks_32_32_32/L2/L2_1_8_2.raw"
    assert G.store_brick_path("/data/vol", 16, 0, (0, 0, 0)) == "/data/vol/bricks_16_16_16/L0/L0_0_0_0.raw"


def _scene(bytes_=2, dims=(40, 36, 28), bs=8, bb=(2, 2, 1), seed=5):
    vc = G.volume_cfg(dims, bs, 1, bytes_, big_brick=bb)
    amp = (60.0, 255.0) if bytes_ == 1 else (3000.0, 60000.0)
    p = G.synth_params(seed, 6, sigma=(2.0, 6.0), amp=amp, noise=4, threads=2)
    return vc, G.Source.synthetic(vc, p)


def _big_bricks(vc):
    for l in range(vc.nb_levels):
        inner = vc.brick_size[0] - 2 * vc.covering[0]
        n = [vc.real_sizes[l][a] // inner // vc.big_brick_size[a] for a in range(3)]
        for z in range(n[2]):
            for y in range(n[1]):
                for x in range(n[0]):
                    yield l, (x, y, z)


@pytest.mark.parametrize("bytes_,bb,compression", [(2, (2, 2, 1), True), (1, (1, 1, 1), True), (2, (1, 2, 2), False)])
def test_store_round_trip(gcv, tmp_path, bytes_, bb, compression):
    vc, src = _scene(bytes_, bb=bb)
    vol = str(tmp_path / "vol") + "/"
    G.store_write(vc, src, vol, compression=compression, threads=4)
    st = G.Source.store(vc, vol)
    n = 0
    for l, b in _big_bricks(vc):
        s0, want = src.get(vc, l, b)
        s1, got = st.get(vc, l, b)
        assert s0 == s1 == 0
        assert np.array_equal(want, got), (l, b)
        path = G.store_brick_path(vol, vc.brick_size[0], l, b)
        raw = open(path, "rb").read()
        if compression:
            assert G.lz4_decompress(raw, want.nbytes) == want.tobytes()
        else:
            assert raw == want.tobytes()
        n += 1
    assert n > 8
    hits, loads, held, empties = st.store_stats()
    assert (hits, loads, held, empties) == (0, n, n, 0)
    for l, b in _big_bricks(vc):
        st.get(vc, l, b)
    assert st.store_stats()[:2] == (n, n)


def test_store_host_lru_bound(gcv, tmp_path):
    vc, src = _scene(1, bb=(1, 1, 1))
    vol = str(tmp_path / "vol")
    G.store_write(vc, src, vol)
    st = G.Source.store(vc, vol, cache_bricks=3)
    keys = list(_big_bricks(vc))
    for l, b in keys[:5]:
        st.get(vc, l, b)
    assert st.store_stats() == (0, 5, 3, 0)
    # the three most recent are held: reading them again hits
    for l, b in keys[2:5]:
        st.get(vc, l, b)
    assert st.store_stats() == (3, 5, 3, 0)
    # the least recently used (the first) was dropped: reading it loads
    st.get(vc, *keys[0])
    assert st.store_stats() == (3, 6, 3, 0)


def test_store_bricks_xml(gcv, tmp_path):
    vc, src = _scene(2, bb=(2, 2, 1))
    thr = 20000.0
    vol = tmp_path / "vol"
    G.store_write(vc, src, str(vol), empty_threshold=thr)
    text = open(vol / "bricks_8_8_8" / "bricks.xml").read()
    assert text.startswith("<Meta>\n    <Levels>\n        <Level level=\"0\" initial_size_X=\"40\"")
    # the configuration reader takes its levels from the written bricks.xml
    cfg = tmp_path / "vol.cfg"
    cfg.write_text('<Fields>\n<Field key="BrickSize" value="8"/>\n<Field key="BigBrickSizeX" value="2"/>\n'
                   '<Field key="BigBrickSizeY" value="2"/>\n<Field key="BigBrickSizeZ" value="1"/>\n'
                   '<Field key="VoxelCovering" value="1"/>\n<Field key="NumberEncodedBytes" value="2"/>\n'
                   '<Field key="NumberChannels" value="1"/>\n<Field key="WorkingDirectory" value="%s/"/>\n'
                   '<Field key="VolumeFile" value="vol.raw"/>\n</Fields>\n' % tmp_path)
    lvc, _ = G.volume_cfg_load(cfg)
    assert lvc.nb_levels == vc.nb_levels
    assert np.array_equal(lvc.initial(), vc.initial()) and np.array_equal(lvc.real(), vc.real())
    # the empty list: exactly the big bricks whose voxels are all below the threshold
    keys, histo = G.store_read_meta(str(vol), 8)
    want = sorted(G.brick_key(l, *b) for l, b in _big_bricks(vc) if (src.get(vc, l, b)[1] < thr).all())
    assert 0 < len(want) < len(list(_big_bricks(vc)))
    assert keys.tolist() == want
    assert len(histo) == 65535 and not histo.any()       # MetaData's default histogram (u16)
    st = G.Source.store(vc, str(vol))
    for l, b in _big_bricks(vc):
        s, _ = st.get(vc, l, b)
        assert s == (1 if G.brick_key(l, *b) in want else 0)
    assert st.store_stats()[3] == len(want)
    # a caller's histogram is written as given
    h = np.arange(256, dtype=np.float32) * 0.5
    vc8, src8 = _scene(1)
    G.store_write(vc8, src8, str(tmp_path / "v8"), histogram=h)
    keys8, h8 = G.store_read_meta(str(tmp_path / "v8"), 8)
    assert len(keys8) == 0 and np.array_equal(h8, h)


def test_store_write_rejects_mismatched_layout(gcv, tmp_path):
    vc, src = _scene(2, bb=(2, 2, 1))
    other = G.volume_cfg((40, 36, 28), 8, 1, 2, big_brick=(1, 1, 1))
    assert G.lib().gc_store_write(C.byref(other), src.h, str(tmp_path).encode(), 1, 0.0, None, 0, 1) == 5
    assert G.lib().gc_source_create_store(C.byref(vc), str(tmp_path / "none").encode(), 0, C.byref(G.vp())) == 4


SHIM_STORE = r'''
#include <tdns/gpucache/CacheManager.hpp>
int main(int, char **argv)
{
    namespace D = tdns::data;
    const D::VolumeConfiguration vc = D::load_volume_configuration(argv[1]);
    D::MetaData md;
    if (!md.load(vc)) { std::printf("no store\n"); return 1; }
    std::printf("%zu %zu\n", md.get_empty_bricks().size(), md.get_histo().size());
    const D::Bkey k = D::get_key(2, tdns::math::Vector3ui(1, 8, 3));
    uint32_t l = 0;
    tdns::math::Vector3ui p;
    D::get_brick_level_position(k, l, p);
    std::printf("%llu %s %u %u %u %u\n", (unsigned long long)k, D::get_brick_name(k).c_str(), l, p[0], p[1], p[2]);
    std::vector<uint16_t> brick((size_t)8 * 8 * 8 * 4);
    const uint32_t bb[3] = {0, 0, 0};
    int status = -1;
    tdns::detail::check(gc_source_get(md.source(), 0, bb, brick.data(), &status));
    unsigned long long sum = 0;
    for (uint16_t v : brick) sum += v;
    std::printf("%d %llu\n", status, sum);
    return 0;
}
'''


def test_cpp_shim_loads_store(gcv, tmp_path):
    """MetaData::load over a written store through the C++ drop-in header:
    empty list, histogram, Bkey helpers and bricks served from the files."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    vc, src = _scene(2, bb=(2, 2, 1))
    G.store_write(vc, src, str(tmp_path / "vol"), empty_threshold=20000.0)
    cfg = tmp_path / "vol.cfg"
    cfg.write_text('<Fields>\n<Field key="BrickSize" value="8"/>\n<Field key="BigBrickSizeX" value="2"/>\n'
                   '<Field key="BigBrickSizeY" value="2"/>\n<Field key="BigBrickSizeZ" value="1"/>\n'
                   '<Field key="VoxelCovering" value="1"/>\n<Field key="NumberEncodedBytes" value="2"/>\n'
                   '<Field key="NumberChannels" value="1"/>\n<Field key="WorkingDirectory" value="%s/"/>\n'
                   '<Field key="VolumeFile" value="vol.raw"/>\n</Fields>\n' % tmp_path)
    prog = tmp_path / "prog.cpp"
    prog.write_text(SHIM_STORE)
    exe = tmp_path / "prog"
    pkg = os.path.dirname(gcv.LIB_PATH)
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(root, "include"),
                           str(prog), "-L" + pkg, "-lgcvol", "-Wl,-rpath," + pkg, "-o", str(exe)])
    out = subprocess.check_output([str(exe), str(cfg)]).decode().splitlines()
    keys, histo = G.store_read_meta(str(tmp_path / "vol"), 8)
    assert out[0] == "%d %d" % (len(keys), len(histo)) and len(keys) > 0
    k = 2 | 1 << 16 | 8 << 32 | 3 << 48
    assert out[1] == "%d L2_1_8_3 2 1 8 3" % k
    st, want = src.get(vc, 0, (0, 0, 0))
    status = 1 if G.brick_key(0, 0, 0, 0) in keys.tolist() else 0
    assert out[2] == "%d %d" % (status, int(want.astype(np.uint64).sum()) if status == 0 else 0)
