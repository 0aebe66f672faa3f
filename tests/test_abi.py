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
