"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Integer and index results (mips, bricks, page directory, LRU, usage and
request buffers, pool contents, render counters) must be bit-exact.  Rendered
RGBA is compared with the FP64-compositing oracle within 1/255 per channel at
both compositing precisions: at precision 64 the device follows the same
sample trace and arithmetic; at precision 32 (the precision the benchmarks
run) it follows the same sample trace with FP32 classification and blending.
Every RGBA comparison's maximum error and RMSE go to the parity log
($GC_PARITY_LOG).
"""
import numpy as np
import pytest

import gcvol as G
import oracle_lib as O
from scenes import Scene, orbit, tf_table, unit_float

pytestmark = pytest.mark.gpu


def _rand_volume(rng, dims, bytes_):
    if bytes_ == 4:
        return rng.random((dims[2], dims[1], dims[0]), dtype=np.float32) * 100.0
    hi = 256 if bytes_ == 1 else 65536
    return rng.integers(0, hi, size=(dims[2], dims[1], dims[0]), dtype=G.DT[bytes_])


@pytest.mark.parametrize("bytes_", [1, 2, 4])
@pytest.mark.parametrize("dims", [(64, 64, 64), (37, 20, 11), (130, 3, 5)])
def test_mip_reduce_bit_exact(gpu, bytes_, dims):
    rng = np.random.default_rng(1000 + bytes_ + dims[0])
    vol = _rand_volume(rng, dims, bytes_)
    expect = O.mip_reduce(vol)
    d_in = G.DeviceBuffer.from_array(vol)
    d_out = G.DeviceBuffer(max(expect.nbytes, 16))
    G.mip_reduce(d_in, dims, bytes_, d_out)
    G.synchronize()
    got = d_out.to_array(vol.dtype, expect.shape)
    assert np.array_equal(got.view(np.uint8), expect.view(np.uint8))


@pytest.mark.parametrize("bytes_,bs,cov", [(1, 16, 1), (2, 8, 1), (1, 8, 0), (2, 32, 2)])
def test_brick_cut_bit_exact(gpu, bytes_, bs, cov):
    rng = np.random.default_rng(7)
    dims = (45, 33, 20)
    vol = _rand_volume(rng, dims, bytes_)
    inner = bs - 2 * cov
    nb = [(d + inner - 1) // inner for d in dims]
    bricks = np.array([[x, y, z] for z in range(nb[2]) for y in range(nb[1]) for x in range(nb[0])], np.uint32)
    d_lv = G.DeviceBuffer.from_array(vol)
    d_b = G.DeviceBuffer.from_array(bricks)
    d_out = G.DeviceBuffer(len(bricks) * bs ** 3 * bytes_)
    G.check(G.lib().gc_brick_cut(d_lv.p, G.ptr(np.asarray(dims, np.uint32)), bytes_, bs, cov, d_b.p,
                                 len(bricks), d_out.p, None))
    G.synchronize()
    got = d_out.to_array(vol.dtype, (len(bricks), bs, bs, bs))
    for i, b in enumerate(bricks):
        exp = O.brick_extract(vol, bs, cov, b)
        assert np.array_equal(got[i], exp), "brick %s" % (b,)


def _compare_state(gc, oc, pool=True):
    assert gc.ts == oc.ts
    assert np.array_equal(gc.mrpd(), oc.mrpd()), "MRPD differs"
    assert np.array_equal(gc.lru(), oc.lru()), "LRU differs"
    assert np.array_equal(gc.usage(), oc.usage()), "usage differs"
    assert np.array_equal(gc.request(), oc.request()), "request buffer differs"
    _check_cell_table(gc)
    if pool:
        assert np.array_equal(gc.pool(), oc.pool()), "pool contents differ"


def _render_pair(scene, gc, oc, cam, prec, W, H, tf, d_tf, frame, otb, ocs, alpha_stop=0.95, tile=0):
    d_st = G.DeviceBuffer.from_array(np.zeros(4, np.uint64))
    view = G.make_view(W, H, cam, alpha_stop=alpha_stop, precision=prec, tile=tile)
    G.render(gc, view, d_tf, len(tf), frame, d_st)
    G.synchronize()
    gst = d_st.to_array(np.uint64, (4,))
    gpx = frame.pixels()
    oview = O.make_view(scene.ocfg, W, H, cam, tf, alpha_stop=alpha_stop)
    opx = np.zeros(W * H, np.uint32)
    ost = oc.render(oview, otb, ocs, opx, precision=64, nthreads=8)
    return gst, G.unpack_rgba(gpx.reshape(-1)), ost, G.unpack_rgba(opx)


@pytest.mark.parametrize("bytes_", [1, 2])
def test_resident_load_and_render(gpu, bytes_, parity_log):
    sc = Scene((64, 64, 64), 8, 1, bytes_)
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    d_levels = sc.device_levels()
    gc.load_all(d_levels)
    oc.load_all()
    _compare_state(gc, oc)
    assert abs(gc.completude() - oc.completude()) == 0
    tf = tf_table()
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 96, 80
    for prec in (64, 32):
        for k in (0, 3):
            frame = G.Frame(W, H, prec)
            otb = np.zeros(W * H, np.float32)
            ocs = np.zeros(W * H * 4)
            gst, g, ost, o = _render_pair(sc, gc, oc, orbit(k), prec, W, H, tf, d_tf, frame, otb, ocs)
            err = np.abs(g - o)
            rec = parity_log(prec, err, frame=k, bytes=bytes_)
            print("prec", prec, "frame", k, "stats", gst.tolist(), "max err", rec["max_err"], "rmse %.4f" % rec["rmse"])
            assert gst[2] == W * H
            assert err.max() <= 1
            # the sample trace (counters, t buffer, usage, requests) is the
            # oracle's at either precision
            assert gst.tolist() == ost.tolist()
            assert np.array_equal(frame.state()[0].reshape(-1), otb)
            assert np.array_equal(gc.usage(), oc.usage())
            assert np.array_equal(gc.request(), oc.request())


@pytest.mark.parametrize("bytes_,tile,prec", [(1, 0, 64), (2, 2, 64), (1, 0, 32), (2, 1, 32), (4, 0, 64), (4, 1, 32)])
def test_out_of_core_sequence(gpu, bytes_, tile, prec, parity_log):
    """Progressive frames with a pool smaller than the volume: every update's
    LRU, page directory, miss list, pool and buffers match the oracle, and
    every launch's pixels are within 1/255 of the FP64 oracle over the same
    cache state (FP32 or FP64 compositing on the device)."""
    sc = Scene((64, 64, 64), 8, 1, bytes_, cache_size=(6, 5, 5), max_requests=50)
    if bytes_ == 4:
        unit_float(sc)
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    tf = tf_table(seed=11)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 64, 48
    frame = G.Frame(W, H, prec)
    errs = []
    otb = np.zeros(W * H, np.float32)
    ocs = np.zeros(W * H * 4)
    steps = 0
    for k in range(3):
        cam = orbit(3 * k)
        frame.reset()
        otb[:] = 0
        ocs[:] = 0
        for it in range(40):
            gst, g, ost, o = _render_pair(sc, gc, oc, cam, prec, W, H, tf, d_tf, frame, otb, ocs, tile=tile)
            assert gst.tolist() == ost.tolist(), "render counters differ at view %d step %d" % (k, it)
            errs.append(np.abs(g - o))
            assert errs[-1].max() <= 1
            assert np.array_equal(frame.state()[0].reshape(-1), otb), "t buffer differs"   # FP32 ray geometry
            gd, od = gc.update(), oc.update()
            assert gd == od
            _compare_state(gc, oc, pool=(it % 5 == 0))
            ga, gne = gc.asked()
            oa, one = oc.asked()
            assert gne == one and np.array_equal(ga, oa)
            steps += 1
            if gd:
                break
    parity_log(prec, np.concatenate(errs), launches=steps, bytes=bytes_)
    assert steps > 3


def _check_cell_table(gc):
    """The ray caster's cell table equals slot << 2 | w derived from the MRPD
    (slot from the entry's pool voxel origin)."""
    m = gc.mrpd().astype(np.uint64)
    bs = int(gc.g.brick_size)
    cx, cy = int(gc.g.cache_size[0]), int(gc.g.cache_size[1])
    slots = m[:, 0] // bs + cx * (m[:, 1] // bs + cy * (m[:, 2] // bs))
    want = np.where(m[:, 3] == 1, slots * 4 + 1, m[:, 3]).astype(np.uint32)
    assert np.array_equal(gc.cell_table(), want), "cell table differs from the page directory"


def _max_mapped_offset(gc, brick_bytes):
    """Largest pool byte offset of a mapped brick (MRPD entry w == 1 holds the
    slot's voxel origin in the 3-D pool)."""
    m = gc.mrpd()
    e = m[m[:, 3] == 1]
    bs = int(gc.g.brick_size)
    cx, cy = int(gc.g.cache_size[0]), int(gc.g.cache_size[1])
    slots = e[:, 0] // bs + cx * (e[:, 1] // bs + cy * (e[:, 2] // bs))
    return int(slots.max()) * brick_bytes


@pytest.mark.parametrize("pool_shape", [(80, 80, 72), (96, 96, 93)])
def test_large_pool_offsets_match_small_pool(gpu, pool_shape):
    """Pools of 3.5 and 6.5 GiB: the first bricks loaded take the last slots
    (victims are the tail of the LRU), at byte offsets whose low 32 bits have
    bit 31 set, below and above 4 GiB.  Frames rendered through them must be
    bit-identical (pixels, t buffer, counters) to the same frames through an
    exactly sized pool, at both compositing precisions."""
    sc = Scene((64, 64, 64), 16, 1, 2, max_requests=64)
    brick_bytes = 16 ** 3 * 2
    small = sc.gpu_cache()
    src = G.Source.pyramid(sc.vc, sc.levels)
    big = G.Cache(sc.vc, G.cache_cfg(pool_shape, 16, max_requests=64, sync=True), src)
    tf = tf_table(seed=17)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 96, 64
    for prec in (32, 64):
        for k in (1, 4):
            out = []
            for c in (small, big):
                f = G.Frame(W, H, prec)
                r = G.render_frame(c, G.make_view(W, H, orbit(k), alpha_stop=0.99, precision=prec), d_tf, len(tf), f)
                tb, cs = f.state()
                out.append((f.pixels().copy(), tb.copy(), cs.copy(), r.as_dict()))
                f.close()
            (p0, t0, c0, r0), (p1, t1, c1, r1) = out
            assert r0["complete"] and r1["complete"]
            for key in ("mapped_samples", "brick_visits", "pixels"):
                assert r0[key] == r1[key], key
            assert np.array_equal(p0, p1), "pixels differ (precision %d, view %d)" % (prec, k)
            assert np.array_equal(t0, t1) and np.array_equal(c0.view(np.uint8), c1.view(np.uint8))
    off = _max_mapped_offset(big, brick_bytes)
    assert (off & 0x80000000) and off >= (1 << 31), "bricks did not reach the intended offsets (%d)" % off
    big.close()
    small.close()


@pytest.mark.parametrize("bytes_,prec,lanes", [(1, 32, 0), (2, 64, 0), (2, 32, 0), (1, 32, 4), (2, 32, 8),
                                               (2, 64, 4)])
def test_ray_list_kernel_matches_lane_kernel(gpu, bytes_, prec, lanes):
    """Launches that resume rays stopped at a miss run either one lane per ray
    or a group of lanes per ray (lanes: 4, 8, or 0 for the list length's
    choice, 16 or 64 here).  The same out-of-core frame sequence rendered
    both ways, on two fresh caches, must give bit-identical pixels, t
    buffers, colour state and render counters, and leave the page directory,
    LRU, usage and request buffers bit-identical after every frame."""
    sc = Scene((96, 80, 64), 8, 1, bytes_, cache_size=(9, 9, 8), max_requests=64)
    caches = [sc.gpu_cache(), sc.gpu_cache()]
    G.set_ray_list_max(caches[0], 0)            # lane per ray throughout
    G.set_ray_list_max(caches[1], 1 << 30)      # every resumed launch a group of lanes per ray
    G.set_ray_group(caches[1], lanes)
    tf = tf_table(seed=23)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 80, 64
    for k in range(4):
        out = []
        for c in caches:
            f = G.Frame(W, H, prec)
            r = G.render_frame(c, G.make_view(W, H, orbit(2 * k, distance=2.2), alpha_stop=0.95 if k % 2 else 0.99,
                                              precision=prec), d_tf, len(tf), f)
            tb, cs = f.state()
            out.append((f.pixels().copy(), tb.copy(), cs.copy(), r.as_dict()))
            f.close()
        (p0, t0, c0, r0), (p1, t1, c1, r1) = out
        print("frame", k, "launches", r0["launches"], "samples", r0["mapped_samples"], flush=True)
        assert r0["complete"] and r1["complete"] and r0["launches"] > 1
        for key in ("launches", "mapped_samples", "brick_visits", "pixels", "unmapped_rays", "bricks_loaded"):
            assert r0[key] == r1[key], (k, key, r0[key], r1[key])
        assert np.array_equal(p0, p1), "pixels differ at frame %d" % k
        assert np.array_equal(t0, t1) and np.array_equal(c0.view(np.uint8), c1.view(np.uint8)), k
        a, b = caches
        assert np.array_equal(a.mrpd(), b.mrpd()) and np.array_equal(a.lru(), b.lru()), k
        assert np.array_equal(a.usage(), b.usage()) and np.array_equal(a.request(), b.request()), k
    for c in caches:
        c.close()


@pytest.mark.parametrize("bytes_,resume", [(1, "lane"), (2, "lane"), (2, "group"), (1, "group4"), (2, "group8")])
def test_brick32_kernel_matches_generic(gpu, bytes_, resume):
    """Bricks of 32 with covering 1 at FP32 run variants of the marches
    compiled for those strides (the tile march with 8x8 tiles, the
    group-per-ray march for resumed rays); 16x4 tiles run the generic tile
    march.  The same out-of-core frame sequence (pool smaller than the
    volume) must give bit-identical pixels, t buffers, colour state, counters
    and cache state either way, after every frame.  resume="lane": both
    caches resume rays with the tile march; "group": the first cache resumes
    them with the group-per-ray march."""
    sc = Scene((160, 128, 96), 32, 1, bytes_, cache_size=(5, 4, 4), max_requests=64)
    caches = [sc.gpu_cache(), sc.gpu_cache()]
    G.set_ray_list_max(caches[0], 1 << 30 if resume.startswith("group") else 0)
    G.set_ray_group(caches[0], int(resume[5:]) if resume[5:] else 0)
    G.set_ray_list_max(caches[1], 0)            # every launch of the second cache is a generic tile march
    tf = tf_table(seed=31)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 96, 80
    for k in range(4):
        out = []
        for c, tile in zip(caches, (0, 1)):
            f = G.Frame(W, H, 32)
            view = G.make_view(W, H, orbit(3 * k, distance=2.0 + 0.3 * k), alpha_stop=0.99, precision=32, tile=tile)
            r = G.render_frame(c, view, d_tf, len(tf), f)
            tb, cs = f.state()
            out.append((f.pixels().copy(), tb.copy(), cs.copy(), r.as_dict()))
            f.close()
        (p0, t0, c0, r0), (p1, t1, c1, r1) = out
        print("brick32 frame", k, "launches", r0["launches"], "samples", r0["mapped_samples"], flush=True)
        assert r0["complete"] and r1["complete"]
        for key in ("launches", "mapped_samples", "brick_visits", "pixels", "unmapped_rays", "bricks_loaded"):
            assert r0[key] == r1[key], (k, key, r0[key], r1[key])
        assert np.array_equal(p0, p1), "pixels differ at frame %d" % k
        assert np.array_equal(t0, t1) and np.array_equal(c0.view(np.uint8), c1.view(np.uint8)), k
        a, b = caches
        assert np.array_equal(a.mrpd(), b.mrpd()) and np.array_equal(a.lru(), b.lru()), k
        assert np.array_equal(a.usage(), b.usage()) and np.array_equal(a.request(), b.request()), k
    for c in caches:
        c.close()


@pytest.mark.parametrize("bytes_", [1, 2])
def test_brick32_fp32_against_oracle(gpu, bytes_, parity_log):
    """The bricks-of-32 variant at FP32 against the FP64 oracle, fully
    resident: RGBA within 1/255, the same sample trace."""
    sc = Scene((96, 96, 80), 32, 1, bytes_)
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    gc.load_all(sc.device_levels())
    oc.load_all()
    _compare_state(gc, oc)
    tf = tf_table(seed=7)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 96, 80
    for k in (0, 2):
        frame = G.Frame(W, H, 32)
        otb = np.zeros(W * H, np.float32)
        ocs = np.zeros(W * H * 4)
        gst, g, ost, o = _render_pair(sc, gc, oc, orbit(k), 32, W, H, tf, d_tf, frame, otb, ocs, alpha_stop=0.99)
        err = np.abs(g - o)
        rec = parity_log(32, err, frame=k, bytes=bytes_)
        print("brick32 fp32 frame", k, "stats", gst.tolist(), "max err", rec["max_err"], "rmse %.4f" % rec["rmse"],
              flush=True)
        assert gst[2] == W * H
        assert err.max() <= 1
        assert gst.tolist() == ost.tolist()
        assert np.array_equal(frame.state()[0].reshape(-1), otb)
        frame.close()
