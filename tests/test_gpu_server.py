"""The serving loop on the GPU (gc_server_*): requests of different sizes,
cameras and transfer functions over several volumes, continuous batching,
admission by brick-pool memory (with evictions), and the plain one-at-a-time
mode.  Each result must be the same alone and in the mix.

Requests resume their stopped rays by exact continuation (gc_view.resume =
1): a ray that stops at a missing brick continues in a later launch at the
stored sample position, so the samples of every ray, and with them its
colour, do not depend on which launches it was marched in.  Every request's
pixels must therefore be bit-identical to the same view rendered alone in one
launch over a fully loaded cache (at FP32 and FP64 compositing), and within
1/255 of the FP64 oracle's one-pass render."""
import threading

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


class OnePass:
    """One-launch renders over fully loaded caches (device) and the FP64
    oracle's one-pass renders, per volume."""

    def __init__(self, scenes):
        self.scenes = scenes
        self.gc, self.oc = {}, {}

    def device(self, vol, w, h, cam, tf, astop, prec):
        sc = self.scenes[vol]
        if vol not in self.gc:
            self.gc[vol] = sc.resident_gpu_cache()
        d_tf = G.DeviceBuffer.from_array(tf)
        f = G.Frame(w, h, prec)
        G.render(self.gc[vol], G.make_view(w, h, cam, alpha_stop=astop, precision=prec), d_tf, len(tf), f)
        px = f.pixels().reshape(-1).copy()
        f.close()
        return px

    def oracle(self, vol, w, h, cam, tf, astop):
        sc = self.scenes[vol]
        if vol not in self.oc:
            self.oc[vol] = sc.oracle_cache()
            self.oc[vol].load_all()
        opx = np.zeros(w * h, np.uint32)
        self.oc[vol].render(O.make_view(sc.ocfg, w, h, cam, tf, alpha_stop=astop), np.zeros(w * h, np.float32),
                            np.zeros(w * h * 4), opx, precision=64, nthreads=8)
        return opx


def _err(a, b):
    return np.abs(G.unpack_rgba(a) - G.unpack_rgba(b))


@pytest.mark.parametrize("plain,budget_vols,prec", [(False, 3, 32), (False, 2, 32), (True, 3, 32), (False, 2, 64)])
def test_server_results_alone_and_in_mix(gpu, plain, budget_vols, prec, parity_log):
    scenes = [Scene(d, 8, 1, b, max_requests=64) for d, b in VOLS]
    pool_bytes = [2048 * 8 ** 3 * b for _, b in VOLS]
    max_views, slot_pixels = 8, 128 * 96
    slot_bytes = max_views * slot_pixels * (4 + 4 + (32 if prec == 64 else 16) + 16 + 8 + 8 + 1)
    budget = sum(sorted(p + slot_bytes for p in pool_bytes)[-budget_vols:]) + 1
    srv = G.Server(budget, max_views=max_views, slot_pixels=slot_pixels, precision=prec, plain=plain)
    srcs = []
    for sc in scenes:
        src = G.Source.pyramid(sc.vc, sc.levels)
        srcs.append(src)
    vids = [srv.add_volume(sc.vc, src, pb, max_requests=64) for sc, src, pb in zip(scenes, srcs, pool_bytes)]
    reqs = _requests(14)
    ids = {}
    print("server plain", plain, "budget volumes", budget_vols, "precision", prec, flush=True)
    for k, (vol, w, h, cam, tseed, astop) in enumerate(reqs):
        view = G.make_view(w, h, cam, alpha_stop=astop, precision=prec, resume=1)
        ids[srv.submit(vids[vol], view, tf_table(seed=tseed))] = k
    srv.drain()
    res = srv.poll(64)
    assert len(res) == len(reqs)
    st = srv.stats()
    assert st.completed == len(reqs) and st.samples > 0
    if budget_vols < len(VOLS):
        assert st.pools_evicted > 0
    got = {}
    multi = 0
    for r in res:
        k = ids[r.id]
        assert (r.width, r.height) == reqs[k][1:3]
        assert r.latency_ms > 0
        multi += r.launches > 1
        got[k] = srv.take_pixels(r.id, r.width, r.height)
    srv.close()
    assert multi > 0, "no request was rendered in more than one launch"
    ref = OnePass(scenes)
    errs = []
    for k, (vol, w, h, cam, tseed, astop) in enumerate(reqs):
        tf = tf_table(seed=tseed)
        alone = ref.device(vol, w, h, cam, tf, astop, prec)
        e = _err(got[k], ref.oracle(vol, w, h, cam, tf, astop))
        errs.append(e.reshape(-1))
        print("request", k, "vol", vol, (w, h), "mix == one-pass alone", bool(np.array_equal(got[k], alone)),
              "| mix vs oracle max", int(e.max()), flush=True)
        assert np.array_equal(got[k], alone), k
        assert e.max() <= 1, k
    parity_log(prec, np.concatenate(errs), requests=len(reqs), plain=plain)


def test_server_over_device_synthesised_volumes(gpu):
    """Volumes whose bricks are synthesised on the GPU serve the same pixels,
    bit for bit, as the host-synthesised volume rendered alone (its own
    progressive loop on a fresh cache)."""
    vols = [((72, 64, 56), 1), ((64, 56, 48), 2)]
    params = [G.synth_params(4242 + i, 16, sigma=(4.0, 16.0), amp=(60.0, 255.0) if b == 1 else (15000.0, 65535.0),
                             noise=8 if b == 1 else 2048, threads=2) for i, (_, b) in enumerate(vols)]
    vcs = [G.volume_cfg(d, 8, 1, b) for d, b in vols]
    pool_bytes = [2048 * 8 ** 3 * b for _, b in vols]
    srv = G.Server(1 << 30, max_views=4, slot_pixels=128 * 96, precision=32)
    srcs = [G.Source.synthetic(vc, p, device=0) for vc, p in zip(vcs, params)]
    vids = [srv.add_volume(vc, src, pb, max_requests=64) for vc, src, pb in zip(vcs, srcs, pool_bytes)]
    reqs = [(k % 2,) + r[1:] for k, r in enumerate(_requests(8, seed=11))]
    ids = {}
    print("server over device-synthesised volumes", flush=True)
    for k, (vol, w, h, cam, tseed, astop) in enumerate(reqs):
        view = G.make_view(w, h, cam, alpha_stop=astop, precision=32, resume=1)
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
        f = G.Frame(w, h, 32)
        rep = G.render_frame(gc, G.make_view(w, h, cam, alpha_stop=astop, precision=32, resume=1), d_tf, len(tf), f)
        assert rep.complete == 1
        e = _err(got[k], f.pixels().reshape(-1))
        print("request", k, "vol", vol, (w, h), "launches alone", rep.launches, "mix vs alone (host bricks) max",
              int(e.max()), flush=True)
        assert e.max() == 0, k
        gc.close()


def test_server_evicts_only_idle_volumes_with_empty_queues(gpu):
    """Pool budget for one volume at a time.  Volume 1 is resident and idle
    when requests for volume 0 (not resident, lower index) and volume 1 are
    queued together: volume 1 must not be evicted for volume 0 while it has
    queued requests (its requests would be placed in slots that the
    eviction then drops).  Every round of submissions drains."""
    scenes = [Scene(d, 8, 1, b, max_requests=64) for d, b in VOLS[:2]]
    pool_bytes = [512 * 8 ** 3 * b for _, b in VOLS[:2]]
    max_views, slot_pixels = 2, 64 * 48
    slot_bytes = max_views * slot_pixels * (4 + 4 + 16 + 16 + 8 + 8 + 1)
    budget = max(p + slot_bytes for p in pool_bytes) + 1
    srv = G.Server(budget, max_views=max_views, slot_pixels=slot_pixels, precision=32)
    srcs = [G.Source.pyramid(sc.vc, sc.levels) for sc in scenes]
    vids = [srv.add_volume(sc.vc, src, pb, max_requests=64) for sc, src, pb in zip(scenes, srcs, pool_bytes)]
    tf = tf_table(seed=5)
    for rnd in range(6):
        view = G.make_view(64, 48, G.camera(10.0, 40.0 * rnd, 2.8), precision=32, resume=1)
        srv.submit(vids[1], view, tf)
        srv.drain()                              # volume 1 resident and idle
        for _ in range(3):                       # queued together: volume 0 first
            srv.submit(vids[0], view, tf)
            srv.submit(vids[1], view, tf)
        finished = threading.Event()
        t = threading.Thread(target=lambda: (srv.drain(), finished.set()), daemon=True)
        t.start()
        assert finished.wait(60.0), "round %d did not drain" % rnd
        for r in srv.poll(64):
            srv.take_pixels(r.id, r.width, r.height)
    st = srv.stats()
    assert st.completed == st.submitted == 6 * 7
    assert st.pools_evicted > 0
    srv.close()


def test_server_fed_by_view_request_messages(gpu):
    """Requests folded from the reference's camera / transfer-function /
    bounding-box messages (libNetwork) render as the same views submitted
    directly: bit-identical pixels to each request rendered alone."""
    vc = G.volume_cfg((72, 64, 56), 8, 1, 1)
    params = G.synth_params(77, 16, sigma=(4.0, 16.0), amp=(60.0, 255.0), noise=8, threads=2)
    pool_bytes = 2048 * 8 ** 3
    srv = G.Server(1 << 30, max_views=4, slot_pixels=128 * 96, precision=32)
    vid = srv.add_volume(vc, G.Source.synthetic(vc, params), pool_bytes, max_requests=64)
    boxes = [((-1, -1, -1), (1, 1, 1)), ((-0.8, -1, -0.6), (1, 0.7, 1))]
    reqs = []
    ids = {}
    for k, (_, w, h, cam, tseed, _) in enumerate(_requests(6, seed=5)):
        r = G.Request()
        for msg in (G.msg_camera(cam), G.msg_transfer_function(tf_table(seed=tseed)),
                    G.msg_bounding_box(*boxes[k % 2])):
            r.apply(msg)
        view = r.view(w, h, precision=32)
        view.resume = 1
        reqs.append((view, r.tf()))
        ids[srv.submit(vid, view, r.tf())] = k
    srv.drain()
    res = srv.poll(64)
    assert len(res) == len(reqs)
    got = {ids[x.id]: srv.take_pixels(x.id, x.width, x.height) for x in res}
    srv.close()
    for k, (view, tf) in enumerate(reqs):
        cam = np.array(list(view.inv_mv), np.float32).reshape(3, 4)
        direct = G.make_view(view.width, view.height, cam, alpha_stop=0.95, precision=32, resume=1, bbox=boxes[k % 2])
        assert bytes(direct) == bytes(view)
        d_tf = G.DeviceBuffer.from_array(tf)
        gc = G.Cache(vc, G.cache_cfg(G.pool_shape(pool_bytes, 8 ** 3), 8, max_requests=64), G.Source.synthetic(vc, params))
        f = G.Frame(view.width, view.height, 32)
        rep = G.render_frame(gc, direct, d_tf, len(tf), f)
        assert rep.complete == 1
        e = _err(got[k], f.pixels().reshape(-1))
        print("message request", k, (view.width, view.height), "max diff", int(e.max()), flush=True)
        assert e.max() == 0, k
        gc.close()
