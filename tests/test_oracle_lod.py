"""Oracle pinning (CPU): the view tables and the level-of-detail rule.

Pins, independently of the oracle's own formulas:
  * the view tables of VolumeRayCaster.cu:136-161 for the benchmark volumes
    (the four of tests/golden/geometry.json and the serving ones): marching
    step tstep = 2/256 in the box [-1,1]^3, level step = half a voxel of the
    level's smallest extent = 1/min(size_l), inverse level sizes 2/size_l;
  * compute_LOD (RayCasters.cu:551-567) by what it means: the coarsest level
    whose voxel spacing, seen along each screen axis, is below the pixel
    footprint on that axis -- for axis-aligned screen axes the spacing is the
    voxel size 2/size_l along that axis -- with anisotropic level sizes and
    footprints, so that taking level k's size for level k+1's, or either axis
    for both, fails;
  * the per-ray sample count and opacity correction on a one-level volume:
    a ray through the box centre takes ceil(2 / step) samples and, with a
    constant transfer-function alpha a, ends at alpha 1 - (1 - a)^(n k) with
    k = step / tstep = 128 / size.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import gcvol as G
import oracle_lib as O
from scenes import Scene

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "geometry.json")


def _configs():
    return json.load(open(GOLDEN))["configs"]


def _cfg_of(dims, bs, cov, nlev):
    initial = [[d >> l for d in dims] for l in range(nlev)]
    return O.make_cfg(initial, initial, bs, cov, 1, (1, 1, 1))


# the serving configuration's six volumes, 32^3 bricks: level count
# ceil(log2(max extent / 32)) + 1 (Mipmapper.cu:443)
SERVING = [(512, 512, 512), (1024, 1024, 1024), (2048, 2048, 2048), (2048, 2048, 1024), (4096, 4096, 4096)]


def _all_configs():
    out = [(c["dims"], c["bs"], c["cov"], c["levels"]) for c in _configs()]
    for d in SERVING:
        out.append((list(d), 32, 1, int(np.ceil(np.log2(max(d) / 32))) + 1))
    return out


@pytest.mark.parametrize("k", range(9))
def test_view_tables_of_configs(oracle, k):
    dims, bs, cov, nlev = _all_configs()[k]
    v = O.make_view(_cfg_of(dims, bs, cov, nlev), 8, 8, np.eye(3, 4, dtype=np.float32),
                    np.zeros((4, 4), np.float32))
    assert v.tstep == np.float32(2.0) / np.float32(256.0) == np.float32(1.0 / 128.0)
    assert v.nb_lods == nlev
    for l in range(nlev):
        size = [d >> l for d in dims]             # power-of-two extents halve exactly
        assert v.lod_step[l] == np.float32(1.0 / min(size)), (k, l)
        for a in range(3):
            assert v.inv_levels_size[l][a] == np.float32(2.0 / size[a]), (k, l, a)


def _lod(view, dleft, dup, pdx, pdy):
    return int(O.lib().or_compute_lod(C.byref(view), O.ptr(np.asarray(dleft, np.float32)),
                                      O.ptr(np.asarray(dup, np.float32)), pdx, pdy))


def _expected_level(sizes, ax_l, ax_u, pdx, pdy):
    """Coarsest level l such that every level 1..l has voxel size below the
    footprint on both screen axes (sizes: level extents; ax_*: the volume
    axis each screen axis runs along)."""
    lev = 0
    for l in range(1, len(sizes)):
        if 2.0 / sizes[l][ax_l] < pdx and 2.0 / sizes[l][ax_u] < pdy:
            lev = l
        else:
            break
    return lev


def test_compute_lod_is_coarsest_level_under_the_footprint(oracle):
    dims, nlev = (256, 64, 128), 6
    sizes = [[d >> l for d in dims] for l in range(nlev)]
    view = O.make_view(_cfg_of(dims, 16, 1, nlev), 8, 8, np.eye(3, 4, dtype=np.float32),
                       np.zeros((4, 4), np.float32))
    axes = np.eye(3, dtype=np.float32)
    switches = sorted({2.0 / s[a] for s in sizes for a in range(3)})
    rng = np.random.default_rng(5)
    pd = np.exp(rng.uniform(np.log(1e-3), np.log(2.0), 4000))
    checked = split = 0
    for ax_l, ax_u in ((0, 1), (1, 2), (2, 0), (1, 0)):
        for i in range(0, len(pd) - 1, 2):
            pdx, pdy = float(np.float32(pd[i])), float(np.float32(pd[i + 1]))
            # stay clear of the switch points (FP32 products there)
            if min(abs(pdx - s) / s for s in switches) < 1e-5 or min(abs(pdy - s) / s for s in switches) < 1e-5:
                continue
            want = _expected_level(sizes, ax_l, ax_u, pdx, pdy)
            got = _lod(view, axes[ax_l], axes[ax_u], pdx, pdy)
            assert got == want, (ax_l, ax_u, pdx, pdy)
            checked += 1
            # cases where one axis alone would allow a coarser level
            one = max(_expected_level(sizes, ax_l, ax_l, pdx, 10.0), _expected_level(sizes, ax_u, ax_u, 10.0, pdy))
            split += one > want
    assert checked > 7000 and split > 500


def test_compute_lod_oblique_axes_use_projected_spacing(oracle):
    """An oblique screen axis sees the length of the voxel's extent vector
    projected through it: for dleft = (1,1,0)/sqrt(2) that is
    sqrt((2/sx)^2 + (2/sy)^2)/sqrt(2)."""
    dims, nlev = (128, 32, 64), 5
    sizes = [[d >> l for d in dims] for l in range(nlev)]
    view = O.make_view(_cfg_of(dims, 16, 1, nlev), 8, 8, np.eye(3, 4, dtype=np.float32),
                       np.zeros((4, 4), np.float32))
    dl = np.array([1, 1, 0], np.float32) / np.float32(np.sqrt(2.0))
    du = np.array([0, 0, 1], np.float32)
    for pdx in np.exp(np.linspace(np.log(0.01), np.log(1.0), 300)):
        pdx = float(np.float32(pdx))
        spacing = [np.hypot(2.0 / s[0], 2.0 / s[1]) / np.sqrt(2.0) for s in sizes]
        if min(abs(pdx - sp) / sp for sp in spacing) < 1e-5:
            continue
        want = 0
        for l in range(1, nlev):
            if spacing[l] < pdx:
                want = l
            else:
                break
        assert _lod(view, dl, du, pdx, 100.0) == want, pdx


def _one_level_scene():
    sc = Scene((16, 16, 16), 16, 1, 1, seed=1, nb_blobs=0)
    assert sc.vc.nb_levels == 1
    sc.levels = [np.full(l.shape, 100, np.uint8) for l in sc.levels]
    return sc


@pytest.mark.parametrize("a", [0.002, 0.03])
def test_sample_count_and_opacity_on_one_level(oracle, a):
    sc = _one_level_scene()
    c = sc.oracle_cache()
    c.load_all()
    tf = np.zeros((256, 4), np.float32)
    tf[:, 3] = np.float32(a)
    tf[:, :3] = 1.0
    # one pixel: the ray through the image centre runs along -z from the eye
    # at z = 3 through the box, from t = 2 to t = 4
    view = O.make_view(sc.ocfg, 1, 1, G.camera(0.0, 0.0, 3.0), tf, alpha_stop=2.0)
    tb = np.zeros(1, np.float32)
    cs = np.zeros(4)
    px = np.zeros(1, np.uint32)
    st = c.render(view, tb, cs, px, precision=64)
    n = int(st[0])
    step = 1.0 / 16.0
    assert abs(n - 2.0 / step) <= 1, n
    k = 128.0 / 16.0                              # step / tstep
    expect = 1.0 - (1.0 - float(np.float32(a))) ** (k * n)
    assert abs(cs[3] - expect) < 1e-12, (cs[3], expect)
    assert abs(float(tb[0]) - (2.0 + n * step)) < 1e-5
