"""Oracle pinning (CPU): mip reduction and bricking with covering.

Pins: independent numpy brute force of floor(sum/8) and of the covering rule
(voxel k of brick b is global b*(bs-2c) + k - c, zero outside), plus the
product's host brick sources checked against the oracle bricker.
"""
import numpy as np
import pytest


def mip_brute(vol):
    z, y, x = (s // 2 for s in vol.shape)
    v = vol[: 2 * z, : 2 * y, : 2 * x]
    if vol.dtype == np.float32:
        acc = np.zeros((z, y, x), np.float64)
        # z, y, x summation order of the reference loop
        for k in range(2):
            for j in range(2):
                for i in range(2):
                    acc += v[k::2, j::2, i::2].astype(np.float64)
        return (acc / 8.0).astype(np.float32)
    s = np.zeros((z, y, x), np.uint64)
    for k in range(2):
        for j in range(2):
            for i in range(2):
                s += v[k::2, j::2, i::2].astype(np.uint64)
    return (s // 8).astype(vol.dtype)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("shape", [(4, 4, 4), (11, 20, 37), (5, 3, 130), (1, 2, 2)])
def test_mip_floor_mean(oracle, dtype, shape):
    rng = np.random.default_rng(sum(shape))
    if dtype == np.float32:
        vol = (rng.random(shape) * 1000).astype(np.float32)
    else:
        vol = rng.integers(0, np.iinfo(dtype).max + 1, size=shape, dtype=dtype)
    got = oracle.mip_reduce(vol)
    exp = mip_brute(vol)
    assert got.shape == exp.shape
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8))


def test_mip_extremes(oracle):
    # all-max voxels stay max (no overflow); floor, not round (7 ones / 8 -> 0)
    v = np.full((2, 2, 2), 255, np.uint8)
    assert oracle.mip_reduce(v)[0, 0, 0] == 255
    v = np.full((2, 2, 2), 65535, np.uint16)
    assert oracle.mip_reduce(v)[0, 0, 0] == 65535
    v = np.ones((2, 2, 2), np.uint8)
    v[0, 0, 0] = 0
    assert oracle.mip_reduce(v)[0, 0, 0] == 0
    v = np.full((2, 2, 2), 3, np.uint16)
    v[1, 1, 1] = 10   # 31 / 8 = 3.875 -> 3
    assert oracle.mip_reduce(v)[0, 0, 0] == 3


def brick_brute(vol, bs, cov, b, bb=(1, 1, 1)):
    inner = bs - 2 * cov
    Z, Y, X = vol.shape
    out = np.zeros((bs * bb[2], bs * bb[1], bs * bb[0]), vol.dtype)
    for sz in range(bb[2]):
        for sy in range(bb[1]):
            for sx in range(bb[0]):
                o = [(b[a] * bb[a] + s) * inner - cov for a, s in zip(range(3), (sx, sy, sz))]
                for k in range(bs):
                    for j in range(bs):
                        for i in range(bs):
                            gx, gy, gz = o[0] + i, o[1] + j, o[2] + k
                            if 0 <= gx < X and 0 <= gy < Y and 0 <= gz < Z:
                                out[sz * bs + k, sy * bs + j, sx * bs + i] = vol[gz, gy, gx]
    return out


@pytest.mark.parametrize("bs,cov,bb", [(8, 1, (1, 1, 1)), (8, 0, (1, 1, 1)), (6, 2, (1, 1, 1)),
                                       (4, 1, (2, 1, 2)), (4, 0, (2, 2, 1))])
def test_brick_covering_rule(oracle, bs, cov, bb):
    rng = np.random.default_rng(bs * 10 + cov)
    vol = rng.integers(1, 256, size=(9, 13, 17), dtype=np.uint8)
    inner = bs - 2 * cov
    nb = [-(-d // inner) for d in (17, 13, 9)]
    nbb = [-(-nb[a] // bb[a]) for a in range(3)]
    for bz in range(nbb[2]):
        for by in range(nbb[1]):
            for bx in range(nbb[0]):
                got = oracle.brick_extract(vol, bs, cov, (bx, by, bz), bb)
                exp = brick_brute(vol, bs, cov, (bx, by, bz), bb)
                assert np.array_equal(got, exp), (bx, by, bz)


def test_brick_interior_is_global_and_border_is_zero(oracle):
    rng = np.random.default_rng(0)
    vol = rng.integers(1, 65536, size=(20, 20, 20), dtype=np.uint16)
    b = oracle.brick_extract(vol, 8, 1, (0, 0, 0))
    # first brick: local 1..7 <-> global 0..6, local 0 is the zero border
    assert np.array_equal(b[1:, 1:, 1:], vol[:7, :7, :7])
    assert (b[0] == 0).all() and (b[:, 0] == 0).all() and (b[:, :, 0] == 0).all()
    b = oracle.brick_extract(vol, 8, 1, (1, 2, 0))
    assert np.array_equal(b[1:, :, :], vol[0:7, 11:19, 5:13])


@pytest.mark.parametrize("bytes_,bb", [(1, (1, 1, 1)), (2, (1, 1, 1)), (1, (2, 2, 1)), (4, (1, 2, 1))])
def test_product_pyramid_source_equals_oracle(oracle, gcv, bytes_, bb):
    G = gcv
    rng = np.random.default_rng(bytes_)
    dims = (37, 25, 18)
    if bytes_ == 4:
        vol = rng.random((dims[2], dims[1], dims[0]), dtype=np.float32)
    else:
        vol = rng.integers(0, 256 ** bytes_, size=(dims[2], dims[1], dims[0]), dtype=G.DT[bytes_])
    vc = G.volume_cfg(dims, 8, 1, bytes_, bb)
    levels = [vol]
    for _ in range(1, vc.nb_levels):
        levels.append(oracle.mip_reduce(levels[-1]))
    src = G.Source.pyramid(vc, levels)
    _, nbb, _ = oracle.init_meta(vc.initial(), 8, 1, bb)
    for l in range(vc.nb_levels):
        for bz in range(nbb[l][2]):
            for by in range(nbb[l][1]):
                for bx in range(nbb[l][0]):
                    st, got = src.get(vc, l, (bx, by, bz))
                    assert st == 0
                    exp = oracle.brick_extract(levels[l], 8, 1, (bx, by, bz), bb)
                    assert np.array_equal(got, exp)


def test_synthetic_source_consistent_with_bricker(oracle, gcv):
    """Level-0 bricks synthesised on demand equal the bricker applied to the
    synthesised whole volume (covering voxels included)."""
    G = gcv
    dims = (70, 40, 33)
    p = G.synth_params(5, 9, sigma=(3.0, 10.0), amp=(30.0, 250.0), noise=6)
    vol = G.synth_volume(p, dims, 1)
    assert vol.max() > 100
    vc = G.volume_cfg(dims, 16, 1, 1)
    src = G.Source.synthetic(vc, p)
    nb, _, _ = oracle.init_meta(vc.initial(), 16, 1)
    for bz in range(nb[0][2]):
        for by in range(nb[0][1]):
            for bx in range(nb[0][0]):
                st, got = src.get(vc, 0, (bx, by, bz))
                assert st == 0
                assert np.array_equal(got, oracle.brick_extract(vol, 16, 1, (bx, by, bz)))


def test_tf_samples_reference_ramps(oracle, gcv):
    """greyRamp (TransferFunction.cpp:34-44): control points (0,0,black) and
    (1,1,white) give sample i = (i/N) in every channel; defaultColorful at a
    control point reproduces that point."""
    grey = np.array([[0, 0, 0, 0, 0], [1, 1, 1, 1, 1]], np.float32)
    for N in (128, 256):
        s = oracle.tf_samples(grey, N)
        x = (np.arange(N, dtype=np.float32) / np.float32(N)).astype(np.float32)
        assert np.allclose(s, np.stack([x, x, x, x], 1), atol=1e-7, rtol=0)
        assert np.array_equal(gcv.tf_samples(grey, N), s)
    colorful = np.array([[0.20, 0.10, 0, 0, 1], [0.35, 0.70, 0, 1, 1], [0.45, 1.0, 0, 1, 0],
                         [0.70, 1.0, 1, 1, 0], [1.0, 1.0, 1, 0, 0]], np.float32)
    s = oracle.tf_samples(colorful, 100)
    # x = 0.45 sits on control point 2 (value 0.45, alpha 1, green)
    assert np.allclose(s[45], [0, 1, 0, 1], atol=1e-6)
    # below the first point the ramp starts at alpha 0 with the first colour
    assert np.allclose(s[0], [0, 0, 1, 0], atol=1e-6)
    assert np.allclose(s[10], [0, 0, 1, 0.05], atol=1e-6)
    assert np.array_equal(gcv.tf_samples(colorful, 100), s)


def test_brick_level_loop_is_per_brick_cut(oracle):
    """or_brick_level (the bricker's per-level loop, used by the CPU
    baseline) cuts brick b of the level at (b % nx, b / nx % ny, b / nx / ny)
    exactly as the per-brick cut does."""
    import oracle_lib as O
    rng = np.random.default_rng(3)
    lv = rng.integers(0, 65536, size=(21, 30, 37), dtype=np.uint16)
    bs, cov = 8, 1
    inner = bs - 2 * cov
    nb = [(d + inner - 1) // inner for d in (37, 30, 21)]
    out = O.brick_level(lv, bs, cov, 3, nb[0] * nb[1] * nb[2])
    for i, b in enumerate(range(3, nb[0] * nb[1] * nb[2])):
        c = (b % nb[0], (b // nb[0]) % nb[1], b // (nb[0] * nb[1]))
        assert np.array_equal(out[i], O.brick_extract(lv, bs, cov, c)), b
