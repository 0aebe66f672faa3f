"""GPU parity of the request-handler paths and of the device page-table API.

* Empty bricks (EmptyThreshold > 0): the handler classifies each big brick,
  asks non-empty bricks first, and the page directory maps Empty bricks to
  (0,0,0,2); the ray caster steps over Empty bricks by tstep
  (RequestHandler.hpp:279-372, CacheManager.hpp:503-632,
  RayCasters.cu:154-159).
* BigBrickSize (2,2,1): requests are deduplicated by big brick and the small
  bricks are copied out of the big-brick staging layout (Bricker_v2.cpp:
  295-299).
  Both run as progressive launches in lockstep with the oracle: after every
  launch pixels within 1/255, counters and t buffers equal; after every
  update the page directory, LRU, usage, requests, pool and the asked list
  (with its non-empty count) bit-identical.
* The device API of include/gc_cache_device.h (gc_get_normalized, gc_get
  with its coarser-level fallback) on a cache whose level 0 is not loaded:
  statuses, values and the usage / request side effects equal the oracle's.
"""
import numpy as np
import pytest

import gcvol as G
import oracle_lib as O
from scenes import Scene, orbit, tf_table
from test_gpu_parity import _compare_state, _render_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bytes_,bb,thr,prec", [(1, (1, 1, 1), 1.0, 64), (2, (2, 2, 1), 0.0, 64),
                                                (1, (2, 2, 1), 1.0, 32), (2, (1, 1, 1), 2000.0, 32)])
def test_empty_and_big_brick_lockstep(gpu, bytes_, bb, thr, prec, parity_log):
    sc = Scene((64, 64, 48), 8, 1, bytes_, cache_size=(6, 5, 5), max_requests=40, big_brick=bb, empty_threshold=thr,
               seed=21, nb_blobs=3, sigma=(3.0, 7.0), noise=0)
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    tf = tf_table(seed=13)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 64, 48
    frame = G.Frame(W, H, prec)
    otb = np.zeros(W * H, np.float32)
    ocs = np.zeros(W * H * 4)
    errs, steps, empties = [], 0, 0
    asked_total = nne_total = 0
    for k in range(3):
        cam = orbit(4 * k, distance=2.4)
        frame.reset()
        otb[:] = 0
        ocs[:] = 0
        for it in range(60):
            print("handler lockstep view", k, "launch", it, flush=True)
            gst, g, ost, o = _render_pair(sc, gc, oc, cam, prec, W, H, tf, d_tf, frame, otb, ocs)
            assert gst.tolist() == ost.tolist(), (k, it)
            errs.append(np.abs(g - o))
            assert errs[-1].max() <= 1, (k, it)
            assert np.array_equal(frame.state()[0].reshape(-1), otb), (k, it)
            gd, od = gc.update(), oc.update()
            assert gd == od
            _compare_state(gc, oc, pool=(it % 4 == 0))
            ga, gne = gc.asked()
            oa, one = oc.asked()
            assert gne == one and np.array_equal(ga, oa), (k, it)
            asked_total += len(oa)
            nne_total += one
            steps += 1
            if gd:
                break
    empties = int((oc.mrpd()[:, 3] == 2).sum())
    print("launches", steps, "asked", asked_total, "non-empty", nne_total, "empty pages", empties, flush=True)
    if thr > 0:
        assert empties > 0 and nne_total < asked_total
    parity_log(prec, np.concatenate(errs), launches=steps, big_brick=list(bb), threshold=thr)


def _map_levels_both(sc, gc, oc, levels):
    g = oc.g
    for _ in range(8):
        req = oc.request()
        ts = oc.ts
        for l in levels:
            for z in range(g.level_dim[l][2]):
                for y in range(g.level_dim[l][1]):
                    for x in range(g.level_dim[l][0]):
                        cell = g.level_coord[l] + x + g.mrpd_dim[0] * (y + g.mrpd_dim[1] * z)
                        if oc.mrpd()[cell, 3] == 0:
                            req[cell] = ts
        gc.write_request(req)
        gd, od = gc.update(), oc.update()
        assert gd == od
        if od:
            return
    raise AssertionError("levels not mapped")


@pytest.mark.parametrize("bytes_,thr", [(1, 0.0), (2, 0.0), (1, 30.0)])
def test_device_api_probes(gpu, bytes_, thr):
    sc = Scene((40, 32, 24), 8, 1, bytes_, seed=3, nb_blobs=4, sigma=(3.0, 8.0), max_requests=10000,
               empty_threshold=thr)
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    _map_levels_both(sc, gc, oc, range(1, sc.vc.nb_levels))
    _compare_state(gc, oc)
    nl = sc.vc.nb_levels
    rng = np.random.default_rng(11 + bytes_)
    n = 3000
    pos = (rng.random((n, 3)) * 0.99).astype(np.float32)
    levels = rng.integers(0, nl, n).astype(np.uint32)
    norm = np.float32(1.0) / np.float32(8.0 * (255.0 if bytes_ == 1 else 65535.0))
    seen = set()
    for mode in (0, 1):
        print("device API probes mode", mode, "bytes", bytes_, "threshold", thr, flush=True)
        vals, sts = gc.probe(mode, pos, levels, init=-1.0)
        for i in range(n):
            st, raw = oc.probe(mode, int(levels[i]), pos[i])
            assert sts[i] == st, (mode, i)
            seen.add((mode, st))
            if st == 2:
                want = np.float32(-1.0)
            elif mode == 0:
                want = np.float32(raw) * norm if st == 0 else np.float32(0.0)
            else:
                want = np.float32(raw)
            assert vals[i] == want, (mode, i, vals[i], want)
        assert np.array_equal(gc.usage(), oc.usage()), mode
        assert np.array_equal(gc.request(), oc.request()), mode
    assert (0, 0) in seen and (0, 1) in seen and (1, 0) in seen
    if thr > 0:
        assert (0, 2) in seen and (1, 2) in seen
