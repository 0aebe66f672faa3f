"""GPU parity: the multi-level page table -- page-table caches between the
MRPD and the brick pool -- through the C ABI against the CPU oracle.

After every update the MRPD, every table cache's entries, block usage and
LRU, the data cache's LRU and usage, the request buffer, the asked list and
the count of unreferenced bricks must be bit-exact; every launch's render
counters and t buffer too, and its RGBA within 1/255 of the FP64 oracle.
The device page-table API (gc_get_normalized / gc_get) walks the tables with
the oracle's answers and side effects.
"""
import numpy as np
import pytest

import gcvol as G
import oracle_lib as O
from scenes import Scene, orbit, tf_table, unit_float

pytestmark = pytest.mark.gpu

# (CacheSize[k+1], BlockSize[k+1]) in configuration order, k = 0 next to the pool
ONE_SMALL = [((3, 2, 2), (2, 2, 2))]
ONE_ODD = [((4, 3, 2), (3, 2, 2))]
TWO = [((4, 3, 3), (2, 2, 2)), ((2, 2, 1), (2, 1, 2))]
THREE = [((4, 3, 3), (2, 2, 1)), ((3, 2, 2), (1, 2, 2)), ((2, 2, 1), (2, 1, 1))]


def _check_cells(gc):
    """The ray caster's MRPD words: table-0 block << 2 | w."""
    m, cells = gc.mrpd(), gc.cell_table()
    t = gc.g.table[0]
    blk = m[:, 0] // t.block[0] + t.blocks[0] * (m[:, 1] // t.block[1] + t.blocks[1] * (m[:, 2] // t.block[2]))
    want = np.where(m[:, 3] == 1, (blk << 2) | 1, m[:, 3])
    assert np.array_equal(cells, want.astype(np.uint32))


def _compare(gc, oc, pool=False):
    assert gc.ts == oc.ts
    assert np.array_equal(gc.mrpd(), oc.mrpd()), "MRPD differs"
    for i in range(oc.nb_tables):
        assert np.array_equal(gc.table(i), oc.table(i)), "table %d entries differ" % i
        assert np.array_equal(gc.table_usage(i), oc.table_usage(i)), "table %d usage differs" % i
        assert np.array_equal(gc.table_lru(i), oc.table_lru(i)), "table %d LRU differs" % i
    assert np.array_equal(gc.lru(), oc.lru()), "LRU differs"
    assert np.array_equal(gc.usage(), oc.usage()), "usage differs"
    assert np.array_equal(gc.request(), oc.request()), "request buffer differs"
    assert gc.unreferenced() == oc.unreferenced()
    _check_cells(gc)
    if pool:
        assert np.array_equal(gc.pool(), oc.pool()), "pool contents differ"


def _render_pair(sc, gc, oc, cam, prec, W, H, tf, d_tf, frame, otb, ocs, compositing=0, tile=0):
    d_st = G.DeviceBuffer.from_array(np.zeros(4, np.uint64))
    view = G.make_view(W, H, cam, precision=prec, tile=tile, compositing=compositing)
    G.render(gc, view, d_tf, len(tf), frame, d_st)
    G.synchronize()
    gst = d_st.to_array(np.uint64, (4,))
    oview = O.make_view(sc.ocfg, W, H, cam, tf)
    opx = np.zeros(W * H, np.uint32)
    if compositing:
        ost = oc.render_mop(oview, otb, ocs, opx, precision=64, nthreads=8)
    else:
        ost = oc.render(oview, otb, ocs, opx, precision=64, nthreads=8)
    return gst, G.unpack_rgba(frame.pixels().reshape(-1)), ost, G.unpack_rgba(opx)


@pytest.mark.parametrize("bytes_,tables,prec,tile,cov", [
    (1, ONE_SMALL, 64, 0, 1), (2, ONE_ODD, 64, 1, 1), (1, TWO, 64, 2, 1), (1, ONE_SMALL, 32, 0, 1),
    (2, TWO, 32, 3, 1), (4, ONE_SMALL, 64, 0, 1), (4, ONE_ODD, 32, 0, 1), (1, ONE_SMALL, 64, 0, 0),
    (1, THREE, 64, 0, 1), (2, THREE, 32, 1, 1)])
def test_out_of_core_sequence_with_tables(gpu, bytes_, tables, prec, tile, cov, parity_log):
    print("tables", tables, "bytes", bytes_, "precision", prec, "tile", tile, "covering", cov, flush=True)
    sc = Scene((64, 64, 64), 8, cov, bytes_, cache_size=(6, 5, 5), max_requests=50, tables=tables)
    if bytes_ == 4:
        unit_float(sc)
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    assert gc.nb_tables == len(tables)
    tf = tf_table(seed=11)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 64, 48
    frame = G.Frame(W, H, prec)
    otb = np.zeros(W * H, np.float32)
    ocs = np.zeros(W * H * 4)
    errs, steps, unref = [], 0, 0
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
            assert np.array_equal(frame.state()[0].reshape(-1), otb), "t buffer differs"
            gd, od = gc.update(), oc.update()
            assert gd == od
            _compare(gc, oc, pool=(it % 5 == 0))
            ga, gne = gc.asked()
            oa, one = oc.asked()
            assert gne == one and np.array_equal(ga, oa)
            unref += oc.unreferenced()
            steps += 1
            if gd:
                break
    rec = parity_log(prec, np.concatenate(errs), launches=steps, bytes=bytes_, tables=len(tables))
    print("launches", steps, "unreferenced bricks", unref, "max err", rec["max_err"], flush=True)
    assert steps > 3


@pytest.mark.parametrize("bytes_", [1, 2])
def test_maximum_projection_with_tables(gpu, bytes_):
    sc = Scene((48, 40, 32), 8, 1, bytes_, cache_size=(5, 4, 4), max_requests=40, tables=ONE_SMALL, seed=4)
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    tf = tf_table(seed=2)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 48, 40
    frame = G.Frame(W, H, 64)
    otb = np.zeros(W * H, np.float32)
    ocs = np.zeros(W * H * 4)
    for it in range(40):
        gst, g, ost, o = _render_pair(sc, gc, oc, orbit(2), 64, W, H, tf, d_tf, frame, otb, ocs, compositing=1)
        assert gst.tolist() == ost.tolist(), it
        assert np.abs(g - o).max() <= 1
        gd, od = gc.update(), oc.update()
        assert gd == od
        _compare(gc, oc)
        if gd:
            break
    assert it > 1


@pytest.mark.parametrize("tables", [ONE_SMALL, TWO, THREE])
def test_random_streams_lockstep(gpu, tables):
    """Random request and usage streams (the loads' inputs) on both caches:
    the load kernels' results equal the oracle's after every update."""
    sc = Scene((64, 48, 40), 8, 1, 1, cache_size=(5, 3, 3), max_requests=23, seed=5, nb_blobs=4,
               sigma=(3.0, 8.0), tables=tables)
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    g = oc.g
    rng = np.random.default_rng(3 + len(tables))
    for step in range(40):
        ts = oc.ts
        req = oc.request()
        for _ in range(rng.integers(0, 30)):
            l = rng.integers(0, sc.vc.nb_levels)
            b = [rng.integers(0, g.level_dim[l][a]) for a in range(3)]
            req[(g.level_coord[l] + b[0]) + g.mrpd_dim[0] * (b[1] + g.mrpd_dim[1] * b[2])] = ts
        u = oc.usage()
        u[rng.random(oc.nslots) < 0.2] = ts
        gc.write_request(req)
        gc.write_usage(u)
        for i in range(oc.nb_tables):
            tu = oc.table_usage(i)
            tu[rng.random(len(tu)) < 0.3] = ts
            gc.write_table_usage(i, tu)
        gd, od = gc.update(), oc.update()
        assert gd == od
        _compare(gc, oc, pool=(step % 8 == 0))


def _request_levels(gc, levels, ts):
    """Every brick cell of the given levels marked requested at ts."""
    g = gc.g
    req = gc.request().reshape(int(g.mrpd_dim[2]), int(g.mrpd_dim[1]), int(g.mrpd_dim[0]))
    for l in levels:
        x0 = int(g.level_coord[l])
        req[:int(g.level_dim[l][2]), :int(g.level_dim[l][1]), x0:x0 + int(g.level_dim[l][0])] = ts
    return req.reshape(-1)


def _map_levels_both(sc, gc, oc, levels):
    """One request of every brick of the levels (the list cap and the pool
    hold them all), loaded by both caches."""
    req = _request_levels(gc, levels, oc.ts)
    gc.write_request(req)
    oc.request()[:] = req
    for _ in range(3):
        gd, od = gc.update(), oc.update()
        assert gd == od
        _compare(gc, oc)
    assert gd


@pytest.mark.parametrize("bytes_,thr", [(1, 0.0), (2, 30.0)])
def test_device_api_probes_with_tables(gpu, bytes_, thr):
    sc = Scene((40, 32, 24), 8, 1, bytes_, seed=3, nb_blobs=4, sigma=(3.0, 8.0), max_requests=10000,
               empty_threshold=thr, tables=[((6, 6, 6), (2, 2, 2))])
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    _map_levels_both(sc, gc, oc, range(1, sc.vc.nb_levels))
    nl = sc.vc.nb_levels
    rng = np.random.default_rng(11 + bytes_)
    n = 2000
    pos = (rng.random((n, 3)) * 0.99).astype(np.float32)
    levels = rng.integers(0, nl, n).astype(np.uint32)
    norm = np.float32(1.0) / np.float32(8.0 * (255.0 if bytes_ == 1 else 65535.0))
    seen = set()
    for mode in (0, 1):
        print("device API probes with tables, mode", mode, "bytes", bytes_, "threshold", thr, flush=True)
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
        assert np.array_equal(gc.table_usage(0), oc.table_usage(0)), mode
    assert (0, 0) in seen and (0, 1) in seen and (1, 0) in seen


def test_resident_tables_match_flat_cache(gpu):
    """A table cache holding every page block, block sizes dividing the bricks
    per axis (big bricks of 2): with every brick loaded, the frames equal
    the MRPD-only cache's bit for bit."""
    kw = dict(dims=(64, 64, 64), bs=8, cov=1, bytes_=1, big_brick=(2, 2, 2), seed=9, max_requests=100000)
    flat = Scene(**kw)
    g0 = O.geometry(Scene(tables=[((1, 1, 1), (2, 2, 2))], **kw).ocfg)
    nblk = sum(int(np.prod(list(g0.top_dim[l]))) for l in range(flat.vc.nb_levels))
    tab = Scene(tables=[((nblk, 1, 1), (2, 2, 2))], **kw)
    caches = []
    for sc in (flat, tab):
        c = sc.gpu_cache()
        c.write_request(_request_levels(c, range(sc.vc.nb_levels), c.ts))
        assert not c.update() and not c.update()
        assert c.unreferenced() == 0
        caches.append(c)
    a, b = caches
    assert np.array_equal(a.lru(), b.lru())
    tf = tf_table(seed=4)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 96, 80
    for prec in (32, 64):
        for k in (0, 5):
            fa, fb = G.Frame(W, H, prec), G.Frame(W, H, prec)
            view = G.make_view(W, H, orbit(k), precision=prec)
            sa = G.DeviceBuffer.from_array(np.zeros(4, np.uint64))
            sb = G.DeviceBuffer.from_array(np.zeros(4, np.uint64))
            G.render(a, view, d_tf, len(tf), fa, sa)
            G.render(b, view, d_tf, len(tf), fb, sb)
            G.synchronize()
            assert np.array_equal(fa.pixels(), fb.pixels()), (prec, k)
            assert np.array_equal(fa.state()[0], fb.state()[0]), (prec, k)
            assert np.array_equal(sa.to_array(np.uint64, (4,)), sb.to_array(np.uint64, (4,))), (prec, k)


def _transparent_pair(max_requests=64, cache_size=(7, 6, 6)):
    """A flat cache and one whose table cache (blocks of 2x2x2 bricks, big
    bricks of 2) holds every page block: the same positions and loads."""
    kw = dict(dims=(64, 64, 64), bs=8, cov=1, bytes_=1, big_brick=(2, 2, 2), seed=12, max_requests=max_requests,
              cache_size=cache_size)
    g0 = O.geometry(Scene(tables=[((1, 1, 1), (2, 2, 2))], **kw).ocfg)
    flat = Scene(**kw)
    nblk = sum(int(np.prod(list(g0.top_dim[l]))) for l in range(flat.vc.nb_levels))
    return flat, Scene(tables=[((nblk, 1, 1), (2, 2, 2))], **kw)


@pytest.mark.parametrize("prec", [32, 64])
def test_frame_driver_with_tables_matches_flat(gpu, prec):
    """gc_render_frame (misses, loads and resumed launches until complete):
    through full transparent tables the frame, its launches and counters
    equal the flat cache's."""
    flat, tab = _transparent_pair()
    a, b = flat.gpu_cache(), tab.gpu_cache()
    tf = tf_table(seed=8)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 80, 64
    for k in (1, 4):
        print("frame driver, precision", prec, "view", k, flush=True)
        fa, fb = G.Frame(W, H, prec), G.Frame(W, H, prec)
        view = G.make_view(W, H, orbit(k), precision=prec)
        ra = G.render_frame(a, view, d_tf, len(tf), fa)
        rb = G.render_frame(b, view, d_tf, len(tf), fb)
        assert ra.complete and rb.complete
        assert ra.launches == rb.launches
        assert ra.stats.mapped_samples == rb.stats.mapped_samples
        assert np.array_equal(fa.pixels(), fb.pixels()), k
        assert np.array_equal(a.lru(), b.lru())
    assert b.unreferenced() == 0


def test_batch_with_tables_matches_flat(gpu):
    flat, tab = _transparent_pair()
    a, b = flat.gpu_cache(), tab.gpu_cache()
    tfs = [tf_table(seed=s) for s in (1, 2, 3)]
    d_tfs = [G.DeviceBuffer.from_array(t) for t in tfs]
    sizes = [len(t) for t in tfs]
    views = [G.make_view(48, 40, orbit(k), precision=32) for k in (0, 2, 5)]
    ba, bb = G.Batch(a, 4, 3 * 48 * 40, 32), G.Batch(b, 4, 3 * 48 * 40, 32)
    ra = ba.render(views, d_tfs, sizes)
    rb = bb.render(views, d_tfs, sizes)
    assert ra.complete and rb.complete and ra.launches == rb.launches
    for i in range(3):
        assert np.array_equal(ba.pixels(i), bb.pixels(i)), i


@pytest.mark.parametrize("bytes_,prec,tables", [(1, 32, ONE_ODD), (2, 64, ONE_SMALL), (1, 32, TWO), (2, 32, THREE)])
def test_ray_list_kernel_with_tables_matches_lane_kernel(gpu, bytes_, prec, tables):
    """Resumed launches through table caches, one lane per ray or one wave
    per ray: bit-identical frames, counters and cache states (tables
    included) after every frame."""
    sc = Scene((96, 80, 64), 8, 1, bytes_, cache_size=(9, 9, 8), max_requests=64, tables=tables)
    caches = [sc.gpu_cache(), sc.gpu_cache()]
    G.set_ray_list_max(caches[0], 0)
    G.set_ray_list_max(caches[1], 1 << 30)
    tf = tf_table(seed=23)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 80, 64
    for k in range(3):
        out = []
        for c in caches:
            f = G.Frame(W, H, prec)
            r = G.render_frame(c, G.make_view(W, H, orbit(2 * k, distance=2.2), alpha_stop=0.95 if k % 2 else 0.99,
                                              precision=prec), d_tf, len(tf), f, max_launches=60)
            tb, cs = f.state()
            out.append((f.pixels().copy(), tb.copy(), cs.copy(), r.as_dict()))
            f.close()
        (p0, t0, c0, r0), (p1, t1, c1, r1) = out
        print("frame", k, "launches", r0["launches"], "lane launches", r1["lane_launches"], "samples",
              r0["mapped_samples"], "complete", r0["complete"], flush=True)
        assert r0["launches"] > 1
        for key in ("launches", "complete", "mapped_samples", "brick_visits", "pixels", "unmapped_rays",
                    "bricks_loaded"):
            assert r0[key] == r1[key], (k, key, r0[key], r1[key])
        assert np.array_equal(p0, p1), "pixels differ at frame %d" % k
        assert np.array_equal(t0, t1) and np.array_equal(c0.view(np.uint8), c1.view(np.uint8)), k
        a, b = caches
        assert np.array_equal(a.mrpd(), b.mrpd()) and np.array_equal(a.lru(), b.lru()), k
        assert np.array_equal(a.usage(), b.usage()) and np.array_equal(a.request(), b.request()), k
        for i in range(len(tables)):
            assert np.array_equal(a.table(i), b.table(i)), (k, i)
            assert np.array_equal(a.table_usage(i), b.table_usage(i)), (k, i)
            assert np.array_equal(a.table_lru(i), b.table_lru(i)), (k, i)
    for c in caches:
        c.close()


@pytest.mark.parametrize("bytes_", [1, 2])
def test_brick32_with_tables_fp32_against_oracle(gpu, bytes_, parity_log):
    """Bricks of 32 with covering 1 through a table cache (the constant-stride
    variants) at FP32, out of core: state bit-exact, counters and t buffer per
    launch, RGBA within 1/255 of the FP64 oracle; then the frame driver's
    resumed lists (wave-per-ray) equal to the lane march."""
    tables = [((4, 4, 2), (2, 2, 2))]
    sc = Scene((160, 128, 96), 32, 1, bytes_, cache_size=(5, 4, 4), max_requests=40, tables=tables, seed=31)
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    tf = tf_table(seed=9)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 96, 80
    frame = G.Frame(W, H, 32)
    otb = np.zeros(W * H, np.float32)
    ocs = np.zeros(W * H * 4)
    errs = []
    for it in range(40):
        gst, g, ost, o = _render_pair(sc, gc, oc, orbit(1), 32, W, H, tf, d_tf, frame, otb, ocs)
        assert gst.tolist() == ost.tolist(), it
        errs.append(np.abs(g - o))
        assert errs[-1].max() <= 1
        assert np.array_equal(frame.state()[0].reshape(-1), otb), it
        gd, od = gc.update(), oc.update()
        assert gd == od
        _compare(gc, oc, pool=(it % 6 == 0))
        if gd:
            break
    rec = parity_log(32, np.concatenate(errs), launches=it + 1, bytes=bytes_, tables=1, brick=32)
    print("brick32 tables launches", it + 1, "max err", rec["max_err"], "rmse %.4f" % rec["rmse"], flush=True)
    assert it > 1
    # resumed lists: lane march vs wave-per-ray march, bit-identical
    caches = [sc.gpu_cache(), sc.gpu_cache()]
    G.set_ray_list_max(caches[0], 0)
    G.set_ray_list_max(caches[1], 1 << 30)
    out = []
    for c in caches:
        f = G.Frame(W, H, 32)
        r = G.render_frame(c, G.make_view(W, H, orbit(3), precision=32), d_tf, len(tf), f, max_launches=60)
        out.append((f.pixels().copy(), f.state()[0].copy(), r.as_dict()))
    assert out[0][2]["launches"] == out[1][2]["launches"] and out[0][2]["mapped_samples"] == out[1][2]["mapped_samples"]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(caches[0].table(0), caches[1].table(0))
    assert np.array_equal(caches[0].table_usage(0), caches[1].table_usage(0))


@pytest.mark.parametrize("bytes_,prec,tables,exited", [(1, 32, ONE_ODD, 0), (2, 64, ONE_SMALL, 0), (2, 32, TWO, 0),
                                                       (1, 32, THREE, 0)])
def test_handoff_with_tables_matches_lane_march(gpu, bytes_, prec, tables, exited):
    """The hand-over through table caches (the walking march stops a ray
    before the visit of a mapped brick, the group march walks, stamps and
    counts it): bit-identical frames, counters and cache states (tables
    included) after every frame, with every ray that has sampled once
    handed over at its next mapped brick (these small launches rarely reach
    a tail trigger)."""
    sc = Scene((96, 80, 64), 8, 1, bytes_, cache_size=(9, 9, 8), max_requests=64, tables=tables)
    caches = [sc.gpu_cache(), sc.gpu_cache()]
    for c in caches:
        G.set_ray_list_max(c, 0)
    G.set_ray_handoff(caches[0], 0)
    G.set_ray_handoff(caches[1], 1, 16, exited)
    log = G.LaunchLog(caches[1], capacity=256)
    tf = tf_table(seed=29)
    d_tf = G.DeviceBuffer.from_array(tf)
    W, H = 80, 64
    handed = 0
    for k in range(3):
        out = []
        for c in caches:
            f = G.Frame(W, H, prec)
            r = G.render_frame(c, G.make_view(W, H, orbit(2 * k + 1, distance=2.2), alpha_stop=0.95 if k % 2 else 0.99,
                                              precision=prec), d_tf, len(tf), f, max_launches=60)
            tb, cs = f.state()
            out.append((f.pixels().copy(), tb.copy(), cs.copy(), r.as_dict()))
            f.close()
        (p0, t0, c0, r0), (p1, t1, c1, r1) = out
        n = sum(r["handed_over"] for r in log.last_frame())
        handed += n
        print("tables handoff frame", k, "launches", r0["launches"], "handed over", n, flush=True)
        for key in ("launches", "complete", "mapped_samples", "brick_visits", "pixels", "unmapped_rays",
                    "bricks_loaded"):
            assert r0[key] == r1[key], (k, key, r0[key], r1[key])
        assert np.array_equal(p0, p1), "pixels differ at frame %d" % k
        assert np.array_equal(t0, t1) and np.array_equal(c0.view(np.uint8), c1.view(np.uint8)), k
        a, b = caches
        assert np.array_equal(a.mrpd(), b.mrpd()) and np.array_equal(a.lru(), b.lru()), k
        assert np.array_equal(a.usage(), b.usage()) and np.array_equal(a.request(), b.request()), k
        for i in range(len(tables)):
            assert np.array_equal(a.table(i), b.table(i)), (k, i)
            assert np.array_equal(a.table_usage(i), b.table_usage(i)), (k, i)
            assert np.array_equal(a.table_lru(i), b.table_lru(i)), (k, i)
    assert handed > 0
    log.close()
    for c in caches:
        c.close()
