"""The C++ drop-in on the GPU: shim_demo (built beside the library from
examples/shim_demo.cpp) follows the reference application's flow through
include/tdns/gpucache/CacheManager.hpp -- .cfg, raw-volume preprocessing on
the GPU, CacheManager<T>, a user kernel sampling through
K_CacheManager::get_normalized (gc_cache_device.h), update() until resident,
one rendered frame.  Its outputs are compared with the oracle:

* every probe's status and value: the value is the box sum of the oracle's
  fully loaded cache at the same (level, position), normalised in FP32
  exactly as the device API does (bit-exact);
* the frame: FP64 compositing against the FP64 oracle caster, <= 1 LSB.
"""
import os
import subprocess

import numpy as np
import pytest

import gcvol as G
import oracle_lib as O
from scenes import Scene, tf_table

pytestmark = pytest.mark.gpu

CFG = ('<Fields>\n<Field key="BrickSize" value="{bs}"/>\n<Field key="BigBrickSizeX" value="1"/>\n'
       '<Field key="BigBrickSizeY" value="1"/>\n<Field key="BigBrickSizeZ" value="1"/>\n'
       '<Field key="VoxelCovering" value="{cov}"/>\n<Field key="NumberEncodedBytes" value="{nb}" />\n'
       '<Field key="NumberChannels" value="1" />\n<Field key="size_X" value="{x}" />\n'
       '<Field key="size_Y" value="{y}" />\n<Field key="size_Z" value="{z}" />\n'
       '<Field key="WorkingDirectory" value="{wd}" />\n<Field key="VolumeFile" value="vol.raw" />\n</Fields>\n')


@pytest.mark.parametrize("bytes_,table_block", [(1, 0), (2, 0), (1, 2), (2, 3)])
def test_shim_demo_matches_oracle(gpu, tmp_path, bytes_, table_block):
    dims = (48, 40, 36)
    tables = ()
    if table_block:
        # CacheSize[1]: a block for every page block (as shim_demo sizes it)
        probe = Scene(dims, 8, 1, bytes_, tables=[((1, 1, 1), (table_block,) * 3)], nb_blobs=0)
        g = O.geometry(probe.ocfg)
        nblk = sum(int(np.prod(list(g.top_dim[l]))) for l in range(probe.vc.nb_levels))
        tables = [((nblk, 1, 1), (table_block,) * 3)]
    sc = Scene(dims, 8, 1, bytes_, tables=tables)
    sc.vol.tofile(tmp_path / "vol.raw")
    cfg = tmp_path / "vol.cfg"
    cfg.write_text(CFG.format(bs=8, cov=1, nb=bytes_, x=dims[0], y=dims[1], z=dims[2], wd=str(tmp_path) + "/"))
    exe = os.path.join(os.path.dirname(G.LIB_PATH), "shim_demo")
    n = 12
    print("shim_demo bytes", bytes_, "dims", dims, "table block", table_block, flush=True)
    out = subprocess.run([exe, str(cfg), str(tmp_path / "out"), str(n), str(table_block)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    fields = out.stdout.split()
    nl = int(fields[1])
    assert nl == sc.vc.nb_levels
    assert fields[fields.index("complete") + 1] == "1"

    vals = np.fromfile(tmp_path / "out.samples", np.float32).reshape(nl, n ** 3)
    st = np.fromfile(tmp_path / "out.status", np.uint32).reshape(nl, n ** 3)
    oc = sc.oracle_cache()
    oc.load_all()
    one = np.float32(1.0)
    norm = one / np.float32(8.0 * (255.0 if bytes_ == 1 else 65535.0))
    fn = np.float32(n)
    i = np.arange(n ** 3)
    px = ((i % n).astype(np.float32) + np.float32(0.5)) / fn
    py = (((i // n) % n).astype(np.float32) + np.float32(0.5)) / fn
    pz = ((i // (n * n)).astype(np.float32) + np.float32(0.5)) / fn
    for lvl in range(nl):
        for k in range(n ** 3):
            ost, s8, _, _ = oc.sample(lvl, (px[k], py[k], pz[k]))
            assert st[lvl, k] == ost, (lvl, k)
            if ost == 0:
                assert vals[lvl, k] == np.float32(s8) * norm, (lvl, k, vals[lvl, k], s8)

    W = H = 128
    tf = tf_table()
    cam = G.camera(20.0, 30.0, 3.0)
    oview = O.make_view(sc.ocfg, W, H, cam, tf, alpha_stop=0.95)
    opx = np.zeros(W * H, np.uint32)
    ost = oc.render(oview, np.zeros(W * H, np.float32), np.zeros(W * H * 4), opx, precision=64, nthreads=8)
    gpx = np.fromfile(tmp_path / "out.pixels", np.uint32)
    err = np.abs(G.unpack_rgba(gpx).astype(np.int32) - G.unpack_rgba(opx).astype(np.int32))
    assert err.max() <= 1
    assert int(fields[fields.index("samples") + 1]) == int(ost[0])
