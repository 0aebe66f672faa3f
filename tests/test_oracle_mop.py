"""Oracle pinning (CPU): the maximum-projection caster (RayCastMOPImpl,
RayCasters.cu:201-349).

* Its march is the DVR march without early termination: on the same cache
  state a MOP launch and a DVR launch with an alpha stop above 1 take the
  same samples, visit the same bricks, stop the same rays at the same t and
  raise the same requests, launch after launch of an out-of-core sequence.
* Its value is the largest sample: on a one-level volume holding a cube of
  value 200 in a background of 10, a ray through the cube's core keeps
  200/255 and a ray that misses the cube keeps 10/255, and the pixel is the
  transfer function at that value with no opacity correction.
"""
import numpy as np
import pytest

import gcvol as G
import oracle_lib as O
from scenes import Scene, orbit, tf_table


def test_mop_trace_is_the_unterminated_dvr_trace(oracle):
    sc = Scene((64, 48, 40), 8, 1, 2, cache_size=(5, 5, 4), max_requests=30, seed=9, nb_blobs=5,
               sigma=(4.0, 12.0), amp=(15000.0, 65535.0), noise=512, empty_threshold=1200.0)
    a, b = sc.oracle_cache(), sc.oracle_cache()
    tf = tf_table(seed=4)
    W, H = 48, 40
    cam = orbit(2, distance=2.4)
    va = O.make_view(sc.ocfg, W, H, cam, tf, alpha_stop=2.0)
    tba, tbb = np.zeros(W * H, np.float32), np.zeros(W * H, np.float32)
    csa, csb = np.zeros(W * H * 4), np.zeros(W * H * 4)
    pa, pb = np.zeros(W * H, np.uint32), np.zeros(W * H, np.uint32)
    launches = 0
    for it in range(80):
        sa = a.render(va, tba, csa, pa, precision=64, nthreads=4)
        sb = b.render_mop(va, tbb, csb, pb, precision=64, nthreads=4)
        assert sa.tolist() == sb.tolist(), it
        assert np.array_equal(tba, tbb), it
        assert np.array_equal(a.request(), b.request()), it
        da, db = a.update(), b.update()
        assert da == db
        assert np.array_equal(a.mrpd(), b.mrpd()) and np.array_equal(a.usage(), b.usage())
        launches += 1
        if da:
            break
    assert da and launches > 3
    assert (a.mrpd()[:, 3] == 2).any()          # Empty pages were stepped over


def _cube_scene(bytes_):
    sc = Scene((16, 16, 16), 16, 1, bytes_, seed=1, nb_blobs=0)
    assert sc.vc.nb_levels == 1
    hi, lo = (200, 10) if bytes_ == 1 else (200 * 257, 10 * 257)
    vol = np.full(sc.levels[0].shape, lo, sc.levels[0].dtype)
    vol[5:11, 5:11, 5:11] = hi
    sc.levels = [vol]
    return sc


def _tf_at(tf, s):
    N = len(tf)
    xf = s * N - 0.5
    fl = np.floor(xf)
    a = xf - fl
    i0 = int(np.clip(fl, 0, N - 1))
    i1 = int(np.clip(fl + 1, 0, N - 1))
    return (1 - a) * tf[i0].astype(np.float64) + a * tf[i1].astype(np.float64)


@pytest.mark.parametrize("bytes_", [1, 2])
def test_mop_keeps_the_largest_sample(oracle, bytes_):
    sc = _cube_scene(bytes_)
    c = sc.oracle_cache()
    c.load_all()
    tf = tf_table(seed=6)
    W = H = 33
    view = O.make_view(sc.ocfg, W, H, G.camera(0.0, 0.0, 3.0), tf, alpha_stop=2.0)
    tb = np.zeros(W * H, np.float32)
    cs = np.zeros(W * H * 4)
    px = np.zeros(W * H, np.uint32)
    st = c.render_mop(view, tb, cs, px, precision=64)
    assert st[3] == 0 and st[0] > 0
    mx = cs.reshape(-1, 4)[:, 0].reshape(H, W)
    # the central ray runs along -z through the cube's core
    assert mx[H // 2, W // 2] == pytest.approx(200 / 255, abs=1e-12)
    # a ray through the volume beside the cube keeps the background; one
    # whose boxes straddle the cube's face keeps half of each (105/255)
    assert mx[H // 2, 6] == pytest.approx(10 / 255, abs=1e-12)
    assert mx[H // 2, 12] == pytest.approx(105 / 255, abs=1e-12)
    assert mx.max() == pytest.approx(200 / 255, abs=1e-12)
    for (yy, xx) in ((H // 2, W // 2), (H // 2, 6), (H // 2, 12)):
        want = np.clip(_tf_at(tf, mx[yy, xx]), 0, 1)
        got = G.unpack_rgba(px.reshape(H, W)[yy, xx:xx + 1])[0]
        assert np.array_equal(got, np.floor(want * 255).astype(got.dtype)), (yy, xx)
