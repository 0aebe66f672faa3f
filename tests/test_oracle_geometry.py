"""Oracle pinning (CPU): geometry, pool shape and the FP32 page-table formulas.

Pins: the hand-derived numbers of tests/golden/geometry.json, closed forms,
and a brute-force round trip of the FP32 brick-position formula.  The
product's host geometry is compared with the oracle's on the same inputs.
"""
import json
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "geometry.json")))


@pytest.mark.parametrize("cfg", GOLD["configs"], ids=lambda c: c["name"])
def test_levels_bricks_directory_pool(oracle, cfg):
    O = oracle
    nl = O.level_count(cfg["dims"], cfg["bs"])
    assert nl == cfg["levels"]
    levels = O.level_sizes(cfg["dims"], nl)
    nb, nbb, real = O.init_meta(levels, cfg["bs"], cfg["cov"])
    assert nb.tolist() == cfg["bricks_per_axis"]
    assert (nbb == nb).all()
    inner = cfg["bs"] - 2 * cfg["cov"]
    assert (real == nb * inner).all()
    if "total_bricks" in cfg:
        assert int(sum(int(np.prod(b)) for b in nb)) == cfg["total_bricks"]
    else:
        assert int(np.prod(nb[0])) == cfg["total_bricks_level0"]
    pool = O.pool_shape(cfg["budget"], cfg["bs"] ** 3 * cfg["bytes"])
    assert list(pool) == cfg["pool"]
    ocfg = O.make_cfg(levels, real, cfg["bs"], cfg["cov"], cfg["bytes"], pool)
    g = O.geometry(ocfg)
    assert list(g.mrpd_dim) == cfg["mrpd_dim"]
    assert int(np.prod(list(g.mrpd_dim))) == cfg["mrpd_cells"]
    for l in range(nl):
        assert list(g.level_dim[l]) == cfg["bricks_per_axis"][l]
        # one pagination level: the virtualised voxels are the real ones and
        # real / virtual is exactly 1 (CacheManager.hpp:186-209)
        assert list(g.virt_vox[l]) == list(real[l])
        assert all(g.rdv[l][a] == 1.0 for a in range(3))
        for a in range(3):
            assert g.ior[l][a] == np.float32(np.float32(levels[l][a]) / np.float32(real[l][a]))
    assert g.level_coord[0] == 0
    for l in range(1, nl):
        assert g.level_coord[l] == g.level_coord[l - 1] + g.level_dim[l - 1][0]


def test_pool_shape_closed_form(oracle):
    # the shape is the smallest round-robin growth of the cube root that holds
    # floor(budget / brick) slots
    for n in [1, 7, 8, 9, 26, 27, 28, 1000, 65536, 1048576, 2432000, 4194304]:
        s = oracle.pool_shape(n * 4096, 4096)
        assert s[0] * s[1] * s[2] >= n
        assert max(s) - min(s) <= 1
        c = int(round(n ** (1 / 3)))
        if c ** 3 == n:
            assert s == (c, c, c)


def test_level_count_matches_log2(oracle):
    for m in [16, 32, 64, 256, 512, 1024, 2048, 4096, 8192]:
        for bs in [8, 16, 32]:
            if m // bs < 1:
                continue
            expect = int(np.ceil(np.log2(m // bs))) + 1
            assert oracle.level_count((m, m // 2 or 1, 3), bs) == expect


def test_fp32_brick_position_round_trip():
    """The page-table position of brick idx, idx/(virt*(1/(bs-2c))) + FLT_EPSILON
    in FP32, maps back to idx under (uint)(pos * bricks) for every level size up
    to 20000 bricks (so the parallel insert never collides and never leaves the
    directory)."""
    eps = np.float32(np.finfo(np.float32).eps)
    for inner in (6, 14, 30, 62, 126, 254, 7, 13):
        one = np.float32(1.0) / np.float32(inner)
        for ldim in range(1, 20001, 7):
            den = np.float32(ldim * inner) * one
            idx = np.arange(ldim, dtype=np.uint32)
            pos = idx.astype(np.float32) / den + eps
            back = (pos * np.float32(ldim)).astype(np.uint32)
            assert np.array_equal(back, idx), (inner, ldim)


def test_oracle_volume_position_matches_formula(oracle):
    levels = oracle.level_sizes((300, 200, 100), 5)
    nb, _, real = oracle.init_meta(levels, 16, 1)
    cfg = oracle.make_cfg(levels, real, 16, 1, 1, (4, 4, 4))
    g = oracle.geometry(cfg)
    import ctypes as C
    rng = np.random.default_rng(3)
    for l in range(5):
        for _ in range(50):
            b = np.array([rng.integers(0, nb[l][a]) for a in range(3)], np.uint32)
            out = np.zeros(3, np.float32)
            oracle.lib().or_volume_position(C.byref(g), l, oracle.ptr(b), oracle.ptr(out))
            one = np.float32(1) / np.float32(14)
            exp = b.astype(np.float32) / (np.array(list(g.virt_vox[l]), np.float32) * one) + np.float32(2 ** -23)
            assert np.array_equal(out, exp.astype(np.float32))
            idx = oracle.lib().or_mrpd_index(C.byref(g), l, oracle.ptr(out))
            x = g.level_coord[l] + b[0]
            assert idx == x + g.mrpd_dim[0] * (b[1] + g.mrpd_dim[1] * b[2])


@pytest.mark.parametrize("cfg", GOLD["configs"], ids=lambda c: c["name"])
def test_product_geometry_equals_oracle(oracle, gcv, cfg):
    G, O = gcv, oracle
    vc = G.volume_cfg(cfg["dims"], cfg["bs"], cfg["cov"], cfg["bytes"])
    assert vc.nb_levels == cfg["levels"]
    cc = G.cache_cfg(cfg["pool"], cfg["bs"])
    g = G.geometry_compute(vc, cc)
    og = O.geometry(O.make_cfg(vc.initial(), vc.real(), cfg["bs"], cfg["cov"], cfg["bytes"], cfg["pool"]))
    for f in ("level_dim", "level_coord", "virt_vox", "mrpd_dim"):
        assert np.array_equal(np.ctypeslib.as_array(getattr(g, f)), np.ctypeslib.as_array(getattr(og, f))), f
    for f in ("ior", "rdv", "one_over"):
        a = np.ctypeslib.as_array(getattr(g, f)).view(np.uint32)
        b = np.ctypeslib.as_array(getattr(og, f)).view(np.uint32)
        assert np.array_equal(a, b), f
    assert g.nslots == int(np.prod(cfg["pool"]))
    assert g.mrpd_cells == cfg["mrpd_cells"]
    assert G.pool_shape(cfg["budget"], cfg["bs"] ** 3 * cfg["bytes"]) == tuple(cfg["pool"])
