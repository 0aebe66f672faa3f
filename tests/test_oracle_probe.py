"""Oracle pinning (CPU): the device page-table API restated (or_probe).

Pins:
  * get_normalized on a mapped brick returns the 2x2x2 texel sum that
    or_sample returns (itself pinned to the box of the level volume by
    test_oracle_dvr.test_sample_is_box_of_global_volume) and stamps usage;
  * get on a mapped brick returns the level volume's voxel at the global
    voxel the FP32 formulas pick (brick * inner + voxel mod inner);
  * with level 0 unmapped, get falls back to the next mapped coarser level,
    returns that level's voxel, and raises a request for the asked level's
    brick only; get_normalized stops at the asked level (Unmapped) and raises
    the same request;
  * Empty bricks answer Empty at the asked level.
"""
import numpy as np

import oracle_lib as O
from scenes import Scene


def _map_levels(c, sc, levels):
    g = c.g
    for _ in range(8):
        req = c.request()
        ts = c.ts
        for l in levels:
            for z in range(g.level_dim[l][2]):
                for y in range(g.level_dim[l][1]):
                    for x in range(g.level_dim[l][0]):
                        cell = g.level_coord[l] + x + g.mrpd_dim[0] * (y + g.mrpd_dim[1] * z)
                        if c.mrpd()[cell, 3] == 0:
                            req[cell] = ts
        if c.update():
            return
    raise AssertionError("levels not mapped")


def _global_voxel(sc, c, l, pos):
    st, s8, b, v = c.sample(l, pos)
    inner = sc.bs - 2 * sc.cov
    gv = b.astype(np.int64) * inner + v.astype(np.int64) % inner
    return st, s8, int(sc.levels[l][gv[2], gv[1], gv[0]])


def test_probe_mapped_and_fallback():
    sc = Scene((40, 32, 24), 8, 1, 1, seed=3, nb_blobs=4, sigma=(3.0, 8.0), max_requests=10000)
    c = sc.oracle_cache()
    _map_levels(c, sc, range(1, sc.vc.nb_levels))
    g = c.g
    rng = np.random.default_rng(2)
    for _ in range(200):
        pos = (rng.random(3) * 0.99).astype(np.float32)
        ts = c.ts
        req0 = c.request().copy()
        # level 1 is mapped: both modes read it
        st1, s8, vox = _global_voxel(sc, c, 1, pos)
        assert st1 == 0
        assert c.probe(0, 1, pos) == (0, float(s8))
        assert c.probe(1, 1, pos) == (0, float(vox))
        # level 0 is not: get falls back to level 1, get_normalized stops
        st, raw = c.probe(1, 0, pos)
        assert (st, raw) == (0, float(vox))
        req = c.request()
        changed = np.nonzero(req != req0)[0]
        # the request cell is the one of the position in the real (padded)
        # level volume, pos * initial/real (K_CacheManager.hpp:365-396)
        real = (pos * np.array(g.ior[0][:], np.float32)).astype(np.float32)
        cell0 = int(O.lib().or_mrpd_index(O.C.byref(g), 0, O.ptr(real)))
        assert set(changed.tolist()) <= {cell0} and req[cell0] == ts
        assert c.probe(0, 0, pos) == (1, 0.0)


def test_probe_empty_bricks():
    sc = Scene((32, 32, 32), 8, 1, 1, seed=1, nb_blobs=0, noise=0, empty_threshold=1.0, max_requests=10000)
    c = sc.oracle_cache()
    _map_levels(c, sc, range(sc.vc.nb_levels))
    rng = np.random.default_rng(4)
    for _ in range(50):
        pos = (rng.random(3) * 0.99).astype(np.float32)
        for mode in (0, 1):
            assert c.probe(mode, 0, pos)[0] == 2
