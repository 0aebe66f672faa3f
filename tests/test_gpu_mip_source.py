"""Mip on miss (gc_source_create_gpu_mip): every brick of every coarse level
built on the GPU from level-0 bricks must be bit-identical to the brick the
CPU preprocessor cuts from its mip level (the oracle's floor(sum/8) mip with
the Mipmapper's accumulator types + the Bricker's covering rule)."""
import itertools

import numpy as np
import pytest

import gcvol as G
from scenes import Scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,bs,cov,bytes_", [((72, 64, 56), 8, 1, 1), ((80, 72, 64), 8, 1, 2),
                                                ((64, 56, 48), 16, 2, 1), ((48, 48, 40), 8, 1, 4)])
def test_gpu_mip_on_miss_bit_exact(gpu, dims, bs, cov, bytes_):
    sc = Scene(dims, bs, cov, bytes_)
    cpu = G.Source.pyramid(sc.vc, sc.levels)
    mip = G.Source.gpu_mip(sc.vc, G.Source.pyramid(sc.vc, sc.levels), memo_bytes=64 << 20)
    inner = bs - 2 * cov
    real = sc.vc.real()
    checked = 0
    print("mip on miss", dims, bs, cov, bytes_, "levels", sc.vc.nb_levels, flush=True)
    for level in range(sc.vc.nb_levels - 1, -1, -1):   # coarsest first: the recursion fills the cache
        nb = [int(real[level][a]) // inner for a in range(3)]
        for bz, by, bx in itertools.product(range(nb[2]), range(nb[1]), range(nb[0])):
            sa, a = cpu.get(sc.vc, level, (bx, by, bz))
            sb, b = mip.get(sc.vc, level, (bx, by, bz))
            assert sa == sb == 0
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (level, bx, by, bz)
            checked += 1
    assert checked > 100


def test_cache_over_gpu_mip_matches_cpu_pyramid(gpu):
    """A cache fed by the GPU mip source evolves exactly like one fed by the
    CPU preprocessor's pyramid: same pages, LRU, usage, pool and pixels after
    every launch and update of an out-of-core sequence."""
    from scenes import tf_table, orbit
    sc = Scene((72, 64, 56), 8, 1, 2, cache_size=(6, 6, 5), max_requests=40)
    ga = sc.gpu_cache()
    lvl0 = G.Source.pyramid(sc.vc, sc.levels)
    mip = G.Source.gpu_mip(sc.vc, lvl0)
    gb = G.Cache(sc.vc, sc.cc, mip)
    tf = tf_table(seed=3)
    d_tf = G.DeviceBuffer.from_array(tf)
    fa, fb = G.Frame(64, 48, 32), G.Frame(64, 48, 32)
    steps = 0
    for k in range(2):
        view = G.make_view(64, 48, orbit(4 * k), precision=32)
        fa.reset()
        fb.reset()
        for it in range(40):
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
