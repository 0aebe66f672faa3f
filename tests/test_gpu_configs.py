"""The benchmark configurations' render paths against the oracle.

* The plain-method configuration at its own size: a 256^3 u8 volume of 12
  seeded blobs plus noise, 16^3 bricks with covering 1, 5 levels, fully
  resident, 512x512 views on the 36-frame orbit (10 degrees per frame),
  alpha stop 0.99, at FP32 and FP64 compositing.
* The out-of-core configuration's path, scaled: u16 voxels, 32^3 bricks with
  covering 1 (the kernels compiled for those strides), a 64-blob volume and a
  pool smaller than it (LRU eviction), FP32 compositing, progressive launches
  (tile march for the first launch, then resumed ray lists with either march)
  in lockstep with the oracle after every launch and every update.
"""
import numpy as np
import pytest

import gcvol as G
import oracle_lib as O
from scenes import Scene, orbit, tf_table
from test_gpu_parity import _compare_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1_scene():
    sc = Scene((256, 256, 256), 16, 1, 1, seed=2024, nb_blobs=12, sigma=(8.0, 40.0), amp=(60.0, 255.0), noise=8)
    assert sc.vc.nb_levels == 5 and sc.total_bricks == 8019
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    gc.load_all(sc.device_levels())
    oc.load_all()
    _compare_state(gc, oc)
    return sc, gc, oc


@pytest.mark.parametrize("prec", [32, 64])
def test_plain_config_full_size(gpu, c1_scene, prec, parity_log):
    sc, gc, oc = c1_scene
    tf = tf_table(seed=1)
    d_tf = G.DeviceBuffer.from_array(tf)
    W = H = 512
    for k in (0, 9, 18, 27):
        cam = orbit(k, rx=20.0, step=10.0, distance=3.0)
        frame = G.Frame(W, H, prec)
        d_st = G.DeviceBuffer.from_array(np.zeros(4, np.uint64))
        G.render(gc, G.make_view(W, H, cam, alpha_stop=0.99, precision=prec), d_tf, len(tf), frame, d_st)
        G.synchronize()
        gst = d_st.to_array(np.uint64, (4,))
        g = G.unpack_rgba(frame.pixels().reshape(-1))
        otb = np.zeros(W * H, np.float32)
        opx = np.zeros(W * H, np.uint32)
        ost = oc.render(O.make_view(sc.ocfg, W, H, cam, tf, alpha_stop=0.99), otb, np.zeros(W * H * 4), opx,
                        precision=64, nthreads=16)
        err = np.abs(g - G.unpack_rgba(opx))
        rec = parity_log(prec, err, frame=k, samples=int(gst[0]))
        print("C1 prec", prec, "frame", k, "samples", int(gst[0]), "max err", rec["max_err"],
              "rmse %.4f" % rec["rmse"], flush=True)
        assert err.max() <= 1
        assert gst.tolist() == ost.tolist()
        assert np.array_equal(frame.state()[0].reshape(-1), otb)
        frame.close()


@pytest.mark.parametrize("ray_list_max,lanes,handoff", [(0, 0, None), (1 << 30, 0, None), (1 << 30, 4, None),
                                                        (0, 0, (1, 8, 0)), (0, 0, (0, 0, 0))])
def test_out_of_core_bricks32_fp32_lockstep(gpu, ray_list_max, lanes, handoff, parity_log):
    """ray_list_max 0: resumed rays run the tile / lane-per-ray march;
    1 << 30: the group-per-ray march (both compiled for bricks of 32), with
    the lanes per ray of the list length or 4.  handoff: the hand-over setting
    of the lane march (None: a new cache's; (1, 8, 0): every ray that has
    sampled once goes to the group march at its next mapped brick; off)."""
    sc = Scene((192, 160, 128), 32, 1, 2, cache_size=(5, 4, 3), max_requests=32, seed=77, nb_blobs=64,
               sigma=(6.0, 30.0), amp=(15000.0, 65535.0), noise=2048)
    assert sc.total_bricks > 3 * 60
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    G.set_ray_list_max(gc, ray_list_max)
    G.set_ray_group(gc, lanes)
    if handoff is not None:
        G.set_ray_handoff(gc, *handoff)
    tf = tf_table(seed=3)
    d_tf = G.DeviceBuffer.from_array(tf)
    sizes = [(128, 96), (96, 72)]
    cams = [G.camera(18.0, 35.0, 2.4), G.camera(-12.0, 200.0, 2.1)]
    views = [G.make_view(w, h, c, alpha_stop=0.99, precision=32) for (w, h), c in zip(sizes, cams)]
    oviews = [O.make_view(sc.ocfg, w, h, c, tf, alpha_stop=0.99) for (w, h), c in zip(sizes, cams)]
    b = G.Batch(gc, 4, 4 * 128 * 96, 32)
    b.begin(views, [d_tf] * 2, [len(tf)] * 2)
    otb = [np.zeros(w * h, np.float32) for w, h in sizes]
    ocs = [np.zeros(w * h * 4) for w, h in sizes]
    prev = np.zeros(4, np.uint64)
    errs = []
    done = False
    for it in range(120):
        print("bricks32 fp32 lockstep launch", it, flush=True)
        b.launch()
        st = b.stats()
        now = np.array([st.mapped_samples, st.brick_visits, st.pixels, st.unmapped_rays], np.uint64)
        delta, prev = now - prev, now
        pend = b.pending()
        ototal = np.zeros(4, np.uint64)
        for i, (w, h) in enumerate(sizes):
            opx = np.zeros(w * h, np.uint32)
            ost = oc.render(oviews[i], otb[i], ocs[i], opx, precision=64, nthreads=16)
            ototal += ost
            err = np.abs(G.unpack_rgba(b.pixels(i)) - G.unpack_rgba(opx))
            errs.append(err)
            assert err.max() <= 1, (it, i)
            tb, cs = b.state(i)
            assert np.array_equal(tb, otb[i]), (it, i)
            assert np.allclose(cs.reshape(-1), ocs[i], rtol=0, atol=2e-5), (it, i)
            assert pend[i] == ost[3], (it, i)
        assert delta[[0, 1, 3]].tolist() == ototal[[0, 1, 3]].tolist(), it
        gd, od = gc.update(), oc.update()
        assert gd == od
        _compare_state(gc, oc, pool=(it % 3 == 0))
        ga, gne = gc.asked()
        oa, one = oc.asked()
        assert gne == one and np.array_equal(ga, oa)
        if gd:
            done = True
            break
    assert done and it > 3
    assert gc.counters()["bricks_loaded"] > 60   # more bricks than slots: eviction happened
    parity_log(32, np.concatenate([e.reshape(-1) for e in errs]), launches=it + 1)
