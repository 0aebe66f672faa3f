"""GPU parity of the maximum-projection caster (RayCastMOPImpl,
RayCasters.cu:201-349) against the FP64 oracle.

* Resident volumes (u8, u16, float), FP32 and FP64: pixels within 1/255,
  the sample / visit / ray counters and the t buffer equal, the stored
  maximum equal (FP64) or within one float rounding (FP32).
* Out of core, progressive: single launches in lockstep with the oracle,
  after every launch pixels within 1/255 and t buffer, counters equal, after
  every update the page directory, LRU, usage and requests bit-identical;
  with exact continuation (resume = 1) the same, its resume positions equal.
* The frame loop (gc_render_frame) with exact continuation ends with the
  one-pass image of the fully loaded volume.
"""
import numpy as np
import pytest

import gcvol as G
import oracle_lib as O
from scenes import Scene, orbit, tf_table
from test_gpu_parity import _compare_state

pytestmark = pytest.mark.gpu


def _mop_pair(sc, gc, oc, cam, prec, W, H, tf, d_tf, frame, otb, ocs, orp=None, resume=0):
    d_st = G.DeviceBuffer.from_array(np.zeros(4, np.uint64))
    view = G.make_view(W, H, cam, precision=prec, compositing=1, resume=resume)
    G.render(gc, view, d_tf, len(tf), frame, d_st)
    G.synchronize()
    gst = d_st.to_array(np.uint64, (4,))
    opx = np.zeros(W * H, np.uint32)
    ost = oc.render_mop(O.make_view(sc.ocfg, W, H, cam, tf), otb, ocs, opx, precision=64, nthreads=8, rpos=orp)
    return gst, G.unpack_rgba(frame.pixels().reshape(-1)), ost, G.unpack_rgba(opx)


@pytest.mark.parametrize("bytes_", [1, 2, 4])
def test_mop_resident(gpu, bytes_, parity_log):
    amp = {1: (60.0, 255.0), 2: (15000.0, 65535.0), 4: (0.2, 1.0)}[bytes_]
    sc = Scene((96, 80, 72), 8, 1, bytes_, seed=31, nb_blobs=8, sigma=(4.0, 14.0), amp=amp, noise=0)
    gc, oc = sc.resident_gpu_cache(), sc.oracle_cache()
    oc.load_all()
    tf = tf_table(seed=8)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 96, 72
    for prec in (32, 64):
        for k in (0, 5):
            frame = G.Frame(W, H, prec)
            otb = np.zeros(W * H, np.float32)
            ocs = np.zeros(W * H * 4)
            print("MOP resident bytes", bytes_, "prec", prec, "view", k, flush=True)
            gst, g, ost, o = _mop_pair(sc, gc, oc, orbit(k, distance=2.6), prec, W, H, tf, d_tf, frame, otb, ocs)
            err = np.abs(g - o)
            parity_log(prec, err, bytes=bytes_, view=k)
            assert err.max() <= 1, (prec, k)
            assert gst.tolist() == ost.tolist() and gst[0] > 0
            tb, cs = frame.state()
            assert np.array_equal(tb.reshape(-1), otb)
            gm = cs.reshape(-1, 4)[:, 0].astype(np.float64)
            om = ocs.reshape(-1, 4)[:, 0]
            if prec == 64:
                assert np.array_equal(gm, om)
            elif bytes_ != 4:
                assert np.array_equal(gm.astype(np.float32), om.astype(np.float32))
            else:
                # float voxels: the FP32 sample (float sum / 8) beside the FP64 one
                assert np.allclose(gm, om, rtol=1e-6, atol=0)
            frame.close()


@pytest.mark.parametrize("prec,resume", [(32, 0), (32, 1), (64, 0)])
def test_mop_out_of_core_lockstep(gpu, prec, resume, parity_log):
    sc = Scene((80, 64, 48), 8, 1, 2, cache_size=(6, 5, 5), max_requests=40, seed=12, nb_blobs=10,
               sigma=(4.0, 16.0), amp=(15000.0, 65535.0), noise=1024, empty_threshold=1500.0)
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    tf = tf_table(seed=2)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 64, 48
    frame = G.Frame(W, H, prec)
    otb = np.zeros(W * H, np.float32)
    ocs = np.zeros(W * H * 4)
    orp = np.zeros(W * H * 4, np.float32) if resume else None
    errs, launches = [], 0
    for k in range(2):
        cam = orbit(6 * k, distance=2.3)
        frame.reset()
        otb[:] = 0
        ocs[:] = 0
        for it in range(80):
            print("MOP out-of-core prec", prec, "resume", resume, "view", k, "launch", it, flush=True)
            gst, g, ost, o = _mop_pair(sc, gc, oc, cam, prec, W, H, tf, d_tf, frame, otb, ocs, orp, resume)
            errs.append(np.abs(g - o))
            assert errs[-1].max() <= 1, (k, it)
            assert gst.tolist() == ost.tolist(), (k, it)
            assert np.array_equal(frame.state()[0].reshape(-1), otb), (k, it)
            gd, od = gc.update(), oc.update()
            assert gd == od
            _compare_state(gc, oc, pool=(it % 4 == 0))
            launches += 1
            if gd:
                break
    assert launches > 6
    assert gc.counters()["bricks_loaded"] > 6 * 5 * 5
    parity_log(prec, np.concatenate(errs), launches=launches, resume=resume)


def test_mop_frame_loop_gives_the_one_pass_image(gpu):
    sc = Scene((80, 64, 48), 8, 1, 1, cache_size=(7, 7, 6), max_requests=25, seed=5, nb_blobs=10,
               sigma=(4.0, 16.0), noise=4)
    gc = sc.gpu_cache()
    full = sc.resident_gpu_cache()
    tf = tf_table(seed=4)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 64, 48
    cam = orbit(3, distance=2.2)
    view = G.make_view(W, H, cam, precision=32, compositing=1, resume=1)
    f = G.Frame(W, H, 32)
    rep = G.render_frame(gc, view, d_tf, len(tf), f)
    print("MOP frame loop", rep.as_dict(), flush=True)
    assert rep.complete and rep.launches > 2
    ref = G.Frame(W, H, 32)
    G.render(full, view, d_tf, len(tf), ref)
    G.synchronize()
    assert np.array_equal(f.pixels(), ref.pixels())
    assert np.array_equal(f.state()[1].reshape(-1, 4)[:, 0], ref.state()[1].reshape(-1, 4)[:, 0])
