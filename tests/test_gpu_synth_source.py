"""Synthetic bricks written on the GPU (gc_source_create_synthetic_gpu): every
brick of every level must be bit-identical to the host synthesiser's brick
(which tests/test_synth.py pins by digest), including covering voxels outside
the volume, uneven level extents, noise and every voxel type; the empty flags
must follow the EmptyThreshold rule; and a cache fed by the device source must
evolve exactly like one fed by the host source."""
import itertools

import numpy as np
import pytest

import gcvol as G
from scenes import orbit, tf_table

pytestmark = pytest.mark.gpu

CASES = [
    # dims, brick, covering, bytes, blobs, noise
    ((70, 50, 33), 16, 1, 1, 12, 8),
    ((96, 80, 64), 32, 1, 2, 64, 2048),
    ((40, 36, 30), 8, 2, 4, 5, 0),
    ((65, 64, 63), 16, 1, 2, 48, 0),
    # a brick size that is not a multiple of 4 (the per-voxel kernel)
    ((50, 41, 30), 10, 1, 1, 7, 8),
]


def _all_keys(vc):
    inner = vc.brick_size[0] - 2 * vc.covering[0]
    real = vc.real()
    keys = []
    for level in range(vc.nb_levels):
        nb = [int(real[level][a]) // inner for a in range(3)]
        for bz, by, bx in itertools.product(range(nb[2]), range(nb[1]), range(nb[0])):
            keys.append((bx, by, bz, level))
    return np.array(keys, np.uint32)


@pytest.mark.parametrize("dims,bs,cov,bytes_,blobs,noise", CASES)
def test_device_bricks_equal_host_bricks(gpu, dims, bs, cov, bytes_, blobs, noise):
    vc = G.volume_cfg(dims, bs, cov, bytes_)
    amp = (60.0, 255.0) if bytes_ == 1 else ((15000.0, 65535.0) if bytes_ == 2 else (0.2, 3.0))
    params = G.synth_params(911 + bs, blobs, sigma=(3.0, 20.0), amp=amp, noise=noise, threads=2)
    host = G.Source.synthetic(vc, params)
    dev = G.Source.synthetic(vc, params, device=0)
    keys = _all_keys(vc)
    print("device synthesis", dims, bs, cov, bytes_, "bricks", len(keys), flush=True)
    got, _ = dev.get_device(vc, keys)
    nonzero = 0
    for i, k in enumerate(keys):
        st, ref = host.get(vc, int(k[3]), (int(k[0]), int(k[1]), int(k[2])))
        assert st == 0
        assert np.array_equal(got[i].view(np.uint8), ref.view(np.uint8)), tuple(k)
        nonzero += int(ref.any())
    assert nonzero > len(keys) // 4


def test_device_empty_flags(gpu):
    vc = G.volume_cfg((96, 80, 64), 16, 1, 1)
    params = G.synth_params(77, 6, sigma=(3.0, 8.0), amp=(60.0, 255.0), noise=8, threads=2)
    dev = G.Source.synthetic(vc, params, device=0)
    keys = _all_keys(vc)
    for thr in (0.0, 20.0, 100.0):
        got, empty = dev.get_device(vc, keys, threshold=thr)
        want = (got.reshape(len(keys), -1).astype(np.float32) < np.float32(thr)).all(axis=1)
        assert np.array_equal(empty.astype(bool), want), thr
        if thr == 20.0:
            assert 0 < want.sum() < len(keys)


@pytest.mark.parametrize("sync", [True, False])
def test_cache_over_device_source_matches_host_source(gpu, sync):
    """Same pages, LRU, usage, pool, asked lists and pixels after every launch
    and update of an out-of-core sequence with eviction, the request handler
    inline or on its worker thread."""
    vc = G.volume_cfg((72, 64, 56), 8, 1, 2)
    params = G.synth_params(5150, 16, sigma=(4.0, 18.0), amp=(15000.0, 65535.0), noise=2048, threads=2)
    cc = G.cache_cfg((6, 6, 5), 8, max_requests=40, sync=sync, empty_threshold=3000.0)
    ga = G.Cache(vc, cc, G.Source.synthetic(vc, params))
    gb = G.Cache(vc, cc, G.Source.synthetic(vc, params, device=0))
    tf = tf_table(seed=3)
    d_tf = G.DeviceBuffer.from_array(tf)
    fa, fb = G.Frame(64, 48, 32), G.Frame(64, 48, 32)
    steps = 0
    for k in range(2):
        view = G.make_view(64, 48, orbit(4 * k), precision=32)
        fa.reset()
        fb.reset()
        for it in range(60):
            print("device source cache sync", sync, "view", k, "launch", it, flush=True)
            G.render(ga, view, d_tf, len(tf), fa)
            G.render(gb, view, d_tf, len(tf), fb)
            G.synchronize()
            assert np.array_equal(fa.pixels(), fb.pixels())
            da, db = ga.update(), gb.update()
            if not sync:
                G.check(G.lib().gc_cache_wait_idle(ga.h))
                G.check(G.lib().gc_cache_wait_idle(gb.h))
            assert da == db
            aa, ab = ga.asked(), gb.asked()
            assert np.array_equal(aa[0], ab[0]) and aa[1] == ab[1]
            assert np.array_equal(ga.mrpd(), gb.mrpd()) and np.array_equal(ga.lru(), gb.lru())
            assert np.array_equal(ga.usage(), gb.usage()) and np.array_equal(ga.pool(), gb.pool())
            steps += 1
            if da:
                break
    assert steps > 3
