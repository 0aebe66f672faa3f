"""A cache fed from the on-disk brick store (gc_source_create_store) evolves
exactly like one fed from the source the store was written from: same pages,
LRU, usage, pool, asked lists and pixels after every launch and update of an
out-of-core sequence with eviction.  The store's empty list (written with the
EmptyThreshold rule) stands in for the cache's own threshold
classification, as the reference's BricksManager empty list does
(BricksManager.cpp:48-77, RequestHandler.hpp:279-372)."""
import numpy as np
import pytest

import gcvol as G
from scenes import orbit, tf_table

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bytes_,bb,thr,compression", [(2, (2, 2, 1), 20000.0, True), (1, (1, 1, 1), 0.0, False)])
def test_cache_over_store_matches_source(gpu, tmp_path, bytes_, bb, thr, compression):
    vc = G.volume_cfg((72, 64, 56), 8, 1, bytes_, big_brick=bb)
    amp = (15000.0, 65535.0) if bytes_ == 2 else (60.0, 255.0)
    params = G.synth_params(4242, 16, sigma=(4.0, 18.0), amp=amp, noise=2048 if bytes_ == 2 else 8, threads=2)
    src = G.Source.synthetic(vc, params)
    vol = str(tmp_path / "vol")
    G.store_write(vc, src, vol, compression=compression, empty_threshold=thr, threads=8)
    keys, _ = G.store_read_meta(vol, 8)
    print("store written: empty big bricks", len(keys), flush=True)
    if thr > 0:
        assert len(keys) > 0
    ga = G.Cache(vc, G.cache_cfg((6, 6, 5), 8, max_requests=40, sync=True, empty_threshold=thr), src)
    store = G.Source.store(vc, vol, cache_bricks=64)
    gb = G.Cache(vc, G.cache_cfg((6, 6, 5), 8, max_requests=40, sync=True, empty_threshold=0.0), store)
    tf = tf_table(seed=3)
    d_tf = G.DeviceBuffer.from_array(tf)
    fa, fb = G.Frame(64, 48, 32), G.Frame(64, 48, 32)
    steps = 0
    for k in range(2):
        view = G.make_view(64, 48, orbit(5 * k), precision=32)
        fa.reset()
        fb.reset()
        for it in range(60):
            print("store cache view", k, "launch", it, flush=True)
            G.render(ga, view, d_tf, len(tf), fa)
            G.render(gb, view, d_tf, len(tf), fb)
            G.synchronize()
            assert np.array_equal(fa.pixels(), fb.pixels())
            da, db = ga.update(), gb.update()
            assert da == db
            aa, ab = ga.asked(), gb.asked()
            assert np.array_equal(aa[0], ab[0]) and aa[1] == ab[1]
            assert np.array_equal(ga.mrpd(), gb.mrpd()) and np.array_equal(ga.lru(), gb.lru())
            assert np.array_equal(ga.usage(), gb.usage()) and np.array_equal(ga.pool(), gb.pool())
            steps += 1
            if da:
                break
    assert steps > 3
    hits, loads, held, empties = store.store_stats()
    print("store stats: hits", hits, "loads", loads, "held", held, "empty answers", empties, flush=True)
    assert loads > 0 and held <= 64
    if thr > 0:
        assert empties > 0
    assert gb.counters()["bricks_loaded"] > 6 * 6 * 5   # more bricks than slots: eviction happened
