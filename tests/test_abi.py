"""The C ABI (CPU): libgcvol.so loads, exports every function include/gcvol.h
declares, the Python binding types every export, and host-only entry points
reject bad input with the documented status codes."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "gcvol.h")


def declared():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gc_[a-z0-9_]+)\s*\(", text)) - {"gc_brick_fn"})


def test_every_declared_symbol_is_exported(gcv):
    names = declared()
    assert len(names) > 40
    out = subprocess.check_output(["nm", "-D", "--defined-only", gcv.LIB_PATH]).decode()
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert set(gcv.EXPORTS) == set(names) - {"gc_brick_fn"}, set(names) ^ set(gcv.EXPORTS)


def test_no_torch_or_oracle_in_product(gcv):
    out = subprocess.check_output(["ldd", gcv.LIB_PATH]).decode()
    assert "oracle" not in out and "torch" not in out
    syms = subprocess.check_output(["nm", "-D", gcv.LIB_PATH]).decode()
    assert "or_render_dvr" not in syms


def test_error_paths(gcv):
    G = gcv
    L = G.lib()
    assert L.gc_volume_cfg_init(None, None, 0, 0, None, 0) == 5          # GC_ERR_ARG
    vc = G.VolumeCfg()
    dims = np.array([64, 64, 64], np.uint32)
    bb = np.array([1, 1, 1], np.uint32)
    assert L.gc_volume_cfg_init(C.byref(vc), G.ptr(dims), 8, 4, G.ptr(bb), 1) == 3   # bs <= 2c
    assert b"VoxelCovering" in L.gc_last_error()
    vc = G.volume_cfg((64, 64, 64), 8, 1, 1)
    cc = G.cache_cfg((4, 4, 4), 16)                                     # BlockSize != BrickSize
    g = G.Geometry()
    assert L.gc_geometry_compute(C.byref(vc), C.byref(cc), C.byref(g)) == 3
    vc.channels = 3
    cc = G.cache_cfg((4, 4, 4), 8)
    assert L.gc_geometry_compute(C.byref(vc), C.byref(cc), C.byref(g)) == 3
    assert L.gc_frame_create(0, 10, 32, C.byref(G.vp())) == 5
    assert L.gc_frame_create(8, 8, 16, C.byref(G.vp())) == 5


def test_cfg_reader(gcv, tmp_path):
    p = tmp_path / "c.cfg"
    p.write_text('<Fields>\n  <Field key="BrickSize" value="16"/>\n  <!-- <Field key="X" value="1"/> -->\n'
                 '  <Field key="VoxelCovering" value="1" />\n</Fields>\n')
    buf = C.create_string_buffer(64)
    assert gcv.lib().gc_cfg_get(str(p).encode(), b"BrickSize", buf, 64) == 0
    assert buf.value == b"16"
    assert gcv.lib().gc_cfg_get(str(p).encode(), b"VoxelCovering", buf, 64) == 0
    assert buf.value == b"1"
    assert gcv.lib().gc_cfg_get(str(p).encode(), b"Missing", buf, 64) == 5
    assert gcv.lib().gc_cfg_get(str(p).encode(), b"X", buf, 64) == 5        # commented out
    # the reference's own sample configuration keys
    ref = ('<Fields>\n<Field key="BrickSize" value="16"/>\n<Field key="BigBrickSizeX" value="1"/>\n'
           '<Field key="NumberEncodedBytes" value="1" />\n<!-- Display -->\n'
           '<Field key="ScreenWidth" value="1920" />\n</Fields>\n')
    p.write_text(ref)
    assert gcv.lib().gc_cfg_get(str(p).encode(), b"ScreenWidth", buf, 64) == 0 and buf.value == b"1920"
    assert gcv.lib().gc_cfg_get(str(p).encode(), b"NumberEncodedBytes", buf, 64) == 0 and buf.value == b"1"


REF_CFG = ('<Fields>\n<Field key="BrickSize" value="16"/>\n<Field key="BigBrickSizeX" value="1"/>\n'
           '<Field key="BigBrickSizeY" value="1"/>\n<Field key="BigBrickSizeZ" value="1"/>\n'
           '<Field key="VoxelCovering" value="1"/>\n<Field key="NumberEncodedBytes" value="1" />\n'
           '<Field key="NumberChannels" value="1" />\n<Field key="MultipleFiles" value="0" />\n'
           '<Field key="size_X" value="256" />\n<Field key="size_Y" value="256" />\n<Field key="size_Z" value="256" />\n'
           '<Field key="WorkingDirectory" value="{wd}" />\n'
           '<Field key="VolumeFile" value="skull_256x256x256_uint8.raw" />\n</Fields>\n')


def test_volume_cfg_load_sizes(gcv, tmp_path):
    """load_volume_configuration on the reference's sample .cfg: without a
    bricks.xml the levels follow the preprocessing rules."""
    p = tmp_path / "skull.cfg"
    p.write_text(REF_CFG.format(wd=str(tmp_path) + "/"))
    vc, path = gcv.volume_cfg_load(p)
    ref = gcv.volume_cfg((256, 256, 256), 16, 1, 1)
    assert vc.nb_levels == ref.nb_levels == 5
    assert np.array_equal(vc.initial(), ref.initial()) and np.array_equal(vc.real(), ref.real())
    assert path == str(tmp_path) + "/skull_256x256x256_uint8.raw"


def test_volume_cfg_load_bricks_xml(gcv, tmp_path):
    """With <WorkingDirectory><base>/bricks_16_16_16/bricks.xml present, the
    level records come from it (VolumeConfiguration.cpp:57-98)."""
    p = tmp_path / "skull.cfg"
    p.write_text(REF_CFG.format(wd=str(tmp_path) + "/"))
    d = tmp_path / "skull_256x256x256_uint8" / "bricks_16_16_16"
    d.mkdir(parents=True)
    (d / "bricks.xml").write_text(
        '<?xml version="1.0"?>\n<Bricks>\n<Levels>\n'
        '<Level initial_size_X="250" initial_size_Y="240" initial_size_Z="230" '
        'real_size_X="252" real_size_Y="252" real_size_Z="238"/>\n'
        '<Level initial_size_X="125" initial_size_Y="120" initial_size_Z="115" '
        'real_size_X="126" real_size_Y="126" real_size_Z="126"/>\n'
        '</Levels>\n<Empty/>\n</Bricks>\n')
    vc, _ = gcv.volume_cfg_load(p)
    assert vc.nb_levels == 2
    assert vc.initial().tolist() == [[250, 240, 230], [125, 120, 115]]
    assert vc.real().tolist() == [[252, 252, 238], [126, 126, 126]]


def test_volume_cfg_load_errors(gcv, tmp_path):
    L = gcv.lib()
    vc = gcv.VolumeCfg()
    assert L.gc_volume_cfg_load(str(tmp_path / "none.cfg").encode(), C.byref(vc), None, 0) == 3
    p = tmp_path / "c.cfg"
    p.write_text(REF_CFG.format(wd="").replace('<Field key="BigBrickSizeY" value="1"/>\n', ""))
    assert L.gc_volume_cfg_load(str(p).encode(), C.byref(vc), None, 0) == 3
    assert b"Missing BigBrickSizeY" in L.gc_last_error()


SHIM_PROG = r'''
#include <tdns/gpucache/CacheManager.hpp>
int main(int, char **argv)
{
    const tdns::data::VolumeConfiguration vc = tdns::data::load_volume_configuration(argv[1]);
    std::printf("%u %u %u %u %u %s %s\n", vc.BrickSize[0], vc.Covering[0], vc.EncodedBytes, vc.NbLevels,
                vc.NbBricks[0][0], vc.VolumeDirectory.c_str(), vc.VolumeFileName.c_str());
    tdns::data::MetaData md;
    tdns::data::CacheConfiguration cc;
    try {
        tdns::gpucache::CacheManager<uint8_t> m(vc, md, cc, 0);
    } catch (const std::runtime_error &e) {
        std::printf("%s\n", e.what());
    }
    try {
        tdns::data::load_volume_configuration("/nonexistent.cfg");
    } catch (const std::runtime_error &e) {
        std::printf("%s\n", e.what());
    }
    return 0;
}
'''


def test_cpp_shim_host_compile_and_config(gcv, tmp_path):
    """The C++ drop-in header compiles with the host compiler alone, links
    against libgcvol.so, and reports errors as std::runtime_error."""
    src = tmp_path / "prog.cpp"
    src.write_text(SHIM_PROG)
    exe = tmp_path / "prog"
    pkg = os.path.dirname(gcv.LIB_PATH)
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
                           str(src), "-L" + pkg, "-lgcvol", "-Wl,-rpath," + pkg, "-o", str(exe)])
    cfg = tmp_path / "skull.cfg"
    cfg.write_text(REF_CFG.format(wd="/data/"))
    out = subprocess.check_output([str(exe), str(cfg)]).decode().splitlines()
    # NbBricks = RealVolumesSizes / BrickSize as the reference computes it
    # (VolumeConfiguration.cpp:101-108): 266 / 16 = 16
    assert out[0] == "16 1 1 5 16 /data/ skull_256x256x256_uint8.raw"
    assert out[1] == "CacheManager: CacheConfiguration needs CacheSize[0] and BlockSize[0]"
    assert out[2].startswith("VolumeConfiguration: unable to open")


def test_shim_demo_built(gcv):
    """The example application over the drop-in headers is built beside the
    library (its run is a GPU test)."""
    exe = os.path.join(os.path.dirname(gcv.LIB_PATH), "shim_demo")
    assert os.access(exe, os.X_OK)
    out = subprocess.check_output(["ldd", exe]).decode()
    assert "libgcvol.so" in out and "oracle" not in out
