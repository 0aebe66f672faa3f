"""The device-resident mip pyramid of a synthetic volume
(gc_source_create_gpu_pyramid): level 0 synthesised on the GPU in z-slabs,
every coarser level reduced on the GPU slab by slab and kept resident.  Every
brick of every level, written on the device (level 0 by synthesis, coarser
levels cut from the resident pyramid), must be bit-identical to the brick the
CPU preprocessor cuts from its own pyramid of the host-synthesised level 0
(the oracle's floor(sum/8) mip with the Mipmapper's accumulator types and the
Bricker's covering rule), whatever the slab thickness, for odd extents and
every voxel type; the EmptyThreshold flags must follow the bricks; and a cache
fed by it must evolve exactly like one fed by the CPU pyramid."""
import itertools

import numpy as np
import pytest

import gcvol as G
from scenes import Scene, orbit, tf_table

pytestmark = pytest.mark.gpu

CASES = [
    # dims, brick, covering, bytes, slab planes
    ((70, 50, 33), 16, 1, 1, 2),
    ((70, 50, 33), 16, 1, 1, 64),
    ((96, 80, 64), 8, 1, 2, 6),
    ((65, 64, 63), 16, 2, 2, 10),
    ((48, 40, 36), 8, 1, 4, 4),
]


def _keys(vc):
    inner = vc.brick_size[0] - 2 * vc.covering[0]
    real = vc.real()
    out = []
    for level in range(vc.nb_levels):
        nb = [int(real[level][a]) // inner for a in range(3)]
        for bz, by, bx in itertools.product(range(nb[2]), range(nb[1]), range(nb[0])):
            out.append((bx, by, bz, level))
    return np.array(out, np.uint32)


def _scene_and_params(dims, bs, cov, bytes_, seed=77):
    amp = (60.0, 255.0) if bytes_ == 1 else ((15000.0, 65535.0) if bytes_ == 2 else (0.2, 3.0))
    kw = dict(nb_blobs=9, sigma=(3.0, 14.0), amp=amp, noise=0 if bytes_ == 4 else 8)
    sc = Scene(dims, bs, cov, bytes_, seed=seed, **kw)
    params = G.synth_params(seed, kw["nb_blobs"], sigma=kw["sigma"], amp=amp, noise=kw["noise"], threads=2)
    return sc, params


@pytest.mark.parametrize("dims,bs,cov,bytes_,slab", CASES)
def test_gpu_pyramid_bricks_bit_exact(gpu, dims, bs, cov, bytes_, slab):
    sc, params = _scene_and_params(dims, bs, cov, bytes_)
    cpu = G.Source.pyramid(sc.vc, sc.levels)
    pyr, ms = G.Source.gpu_pyramid(sc.vc, params, slab_planes=slab)
    keys = _keys(sc.vc)
    print("gpu pyramid", dims, bs, cov, bytes_, "slab", slab, "levels", sc.vc.nb_levels, "bricks", len(keys),
          "build %.1f ms" % ms, flush=True)
    thr = {1: 40.0, 2: 9000.0, 4: 0.5}[bytes_]
    got, empty = pyr.get_device(sc.vc, keys, threshold=thr)
    for i, k in enumerate(keys):
        st, want = cpu.get(sc.vc, int(k[3]), (int(k[0]), int(k[1]), int(k[2])))
        assert st == 0
        assert np.array_equal(got[i].view(np.uint8), want.view(np.uint8)), tuple(k)
        assert bool(empty[i]) == bool((want.astype(np.float32) < np.float32(thr)).all()), tuple(k)
    # host bricks of the coarse levels come from the same pyramid
    for k in keys[-3:]:
        _, a = pyr.get(sc.vc, int(k[3]), (int(k[0]), int(k[1]), int(k[2])))
        _, b = cpu.get(sc.vc, int(k[3]), (int(k[0]), int(k[1]), int(k[2])))
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_cache_over_gpu_pyramid_matches_cpu_pyramid(gpu):
    sc, params = _scene_and_params((72, 64, 56), 8, 1, 2, seed=31)
    sc = Scene((72, 64, 56), 8, 1, 2, cache_size=(6, 6, 5), max_requests=40, seed=31, nb_blobs=9,
               sigma=(3.0, 14.0), amp=(15000.0, 65535.0), noise=8)
    ga = sc.gpu_cache()
    pyr, _ = G.Source.gpu_pyramid(sc.vc, params, slab_planes=8)
    gb = G.Cache(sc.vc, sc.cc, pyr)
    tf = tf_table(seed=3)
    d_tf = G.DeviceBuffer.from_array(tf)
    fa, fb = G.Frame(64, 48, 32), G.Frame(64, 48, 32)
    steps = 0
    for k in range(2):
        view = G.make_view(64, 48, orbit(4 * k), precision=32)
        fa.reset()
        fb.reset()
        for it in range(40):
            print("cache over gpu pyramid view", k, "launch", it, flush=True)
            G.render(ga, view, d_tf, len(tf), fa)
            G.render(gb, view, d_tf, len(tf), fb)
            G.synchronize()
            assert np.array_equal(fa.pixels(), fb.pixels())
            da, db = ga.update(), gb.update()
            assert da == db
            assert np.array_equal(ga.mrpd(), gb.mrpd()) and np.array_equal(ga.lru(), gb.lru())
            assert np.array_equal(ga.usage(), gb.usage()) and np.array_equal(ga.pool(), gb.pool())
            steps += 1
            if da:
                break
    assert steps > 3
