"""The serving loop on the GPU (gc_server_*): requests of different sizes,
cameras and transfer functions over several volumes, continuous batching,
admission by brick-pool memory (with evictions), and the plain one-at-a-time
mode.  Each result must be the same alone and in the mix: every request is
also rendered alone (gc_render_frame on a fresh cache of its volume) and by
the FP64 oracle over a fully loaded cache.

Tolerance.  A ray that stops at a missing brick is resumed by the next
launch from t with pos = origin + dir * t (RayCasters.cu:85-120), so a
multi-launch render may differ from a one-pass render by a few LSB on the
pixels whose rays were resumed at different points.  Alone and in the mix,
bricks arrive at different times, so rays resume at different points: the
comparison allows 4 LSB on at most 0.5 % of a request's channels and 1 LSB
elsewhere, against the alone render and against the oracle."""
import numpy as np
import pytest

import gcvol as G
import oracle_lib as O
from scenes import Scene, tf_table

pytestmark = pytest.mark.gpu

VOLS = [((64, 64, 64), 1), ((80, 72, 64), 2), ((48, 56, 40), 1)]
SIZES = [(64, 48), (96, 80), (40, 40), (128, 72), (72, 96), (33, 17)]


def _requests(n, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        vol = int(rng.integers(len(VOLS)))
        w, h = SIZES[int(rng.integers(len(SIZES)))]
        cam = G.camera(float(rng.uniform(-30, 30)), float(rng.uniform(0, 360)), float(rng.uniform(2.4, 3.4)))
        out.append((vol, w, h, cam, 5 + i % 4, 0.95 if i % 2 else 0.99))
    return out


def _within(a, b):
    err = np.abs(G.unpack_rgba(a) - G.unpack_rgba(b))
    return err.max() <= 4 and (err > 1).mean() <= 0.005, int(err.max()), float((err > 1).mean())


@pytest.mark.parametrize("plain,budget_vols", [(False, 3), (False, 2), (True, 3)])
def test_server_results_alone_and_in_mix(gpu, plain, budget_vols):
    scenes = [Scene(d, 8, 1, b, max_requests=64) for d, b in VOLS]
    pool_bytes = [2048 * 8 ** 3 * b for _, b in VOLS]
    max_views, slot_pixels = 8, 128 * 96
    slot_bytes = max_views * slot_pixels * (4 + 4 + 32 + 8)
    budget = sum(sorted(p + slot_bytes for p in pool_bytes)[-budget_vols:]) + 1
    srv = G.Server(budget, max_views=max_views, slot_pixels=slot_pixels, precision=64, plain=plain)
    srcs = []
    for sc in scenes:
        src = G.Source.pyramid(sc.vc, sc.levels)
        srcs.append(src)
    vids = [srv.add_volume(sc.vc, src, pb, max_requests=64) for sc, src, pb in zip(scenes, srcs, pool_bytes)]
    reqs = _requests(14)
    ids = {}
    print("server plain", plain, "budget volumes", budget_vols, flush=True)
    for k, (vol, w, h, cam, tseed, astop) in enumerate(reqs):
        view = G.make_view(w, h, cam, alpha_stop=astop, precision=64)
        ids[srv.submit(vids[vol], view, tf_table(seed=tseed))] = k
    srv.drain()
    res = srv.poll(64)
    assert len(res) == len(reqs)
    st = srv.stats()
    assert st.completed == len(reqs) and st.samples > 0
    if budget_vols < len(VOLS):
        assert st.pools_evicted > 0
    got = {}
    for r in res:
        k = ids[r.id]
        assert (r.width, r.height) == reqs[k][1:3]
        assert r.latency_ms > 0
        got[k] = srv.take_pixels(r.id, r.width, r.height)
    srv.close()
    ocs = {}
    for k, (vol, w, h, cam, tseed, astop) in enumerate(reqs):
        sc = scenes[vol]
        tf = tf_table(seed=tseed)
        d_tf = G.DeviceBuffer.from_array(tf)
        gc = sc.gpu_cache()
        f = G.Frame(w, h, 64)
        rep = G.render_frame(gc, G.make_view(w, h, cam, alpha_stop=astop, precision=64), d_tf, len(tf), f)
        assert rep.complete == 1
        alone = f.pixels().reshape(-1)
        if vol not in ocs:
            ocs[vol] = sc.oracle_cache()
            ocs[vol].load_all()
        opx = np.zeros(w * h, np.uint32)
        ocs[vol].render(O.make_view(sc.ocfg, w, h, cam, tf, alpha_stop=astop), np.zeros(w * h, np.float32),
                        np.zeros(w * h * 4), opx, precision=64, nthreads=8)
        ok1, e1, f1 = _within(got[k], alone)
        ok2, e2, f2 = _within(got[k], opx)
        print("request", k, "vol", vol, (w, h), "mix vs alone max", e1, "frac>1", round(f1, 5),
              "| mix vs oracle max", e2, "frac>1", round(f2, 5), flush=True)
        assert ok1 and ok2, k


def test_server_over_device_synthesised_volumes(gpu):
    """Volumes whose bricks are synthesised on the GPU serve the same pixels
    as the host-synthesised volume rendered alone."""
    vols = [((72, 64, 56), 1), ((64, 56, 48), 2)]
    params = [G.synth_params(4242 + i, 16, sigma=(4.0, 16.0), amp=(60.0, 255.0) if b == 1 else (15000.0, 65535.0),
                             noise=8 if b == 1 else 2048, threads=2) for i, (_, b) in enumerate(vols)]
    vcs = [G.volume_cfg(d, 8, 1, b) for d, b in vols]
    pool_bytes = [2048 * 8 ** 3 * b for _, b in vols]
    srv = G.Server(1 << 30, max_views=4, slot_pixels=128 * 96, precision=64)
    srcs = [G.Source.synthetic(vc, p, device=0) for vc, p in zip(vcs, params)]
    vids = [srv.add_volume(vc, src, pb, max_requests=64) for vc, src, pb in zip(vcs, srcs, pool_bytes)]
    reqs = [(k % 2,) + r[1:] for k, r in enumerate(_requests(8, seed=11))]
    ids = {}
    print("server over device-synthesised volumes", flush=True)
    for k, (vol, w, h, cam, tseed, astop) in enumerate(reqs):
        view = G.make_view(w, h, cam, alpha_stop=astop, precision=64)
        ids[srv.submit(vids[vol], view, tf_table(seed=tseed))] = k
    srv.drain()
    res = srv.poll(64)
    assert len(res) == len(reqs)
    got = {ids[r.id]: srv.take_pixels(r.id, r.width, r.height) for r in res}
    srv.close()
    for k, (vol, w, h, cam, tseed, astop) in enumerate(reqs):
        tf = tf_table(seed=tseed)
        d_tf = G.DeviceBuffer.from_array(tf)
        cc = G.cache_cfg(G.pool_shape(pool_bytes[vol], 8 ** 3 * vols[vol][1]), 8, max_requests=64)
        gc = G.Cache(vcs[vol], cc, G.Source.synthetic(vcs[vol], params[vol]))
        f = G.Frame(w, h, 64)
        rep = G.render_frame(gc, G.make_view(w, h, cam, alpha_stop=astop, precision=64), d_tf, len(tf), f)
        assert rep.complete == 1
        ok, e, fr = _within(got[k], f.pixels().reshape(-1))
        print("request", k, "vol", vol, (w, h), "mix vs alone (host bricks) max", e, "frac>1", round(fr, 5), flush=True)
        assert ok, k
        gc.close()
