"""Batched rendering on the GPU (gc_batch_*): many uneven views of one cache
in each launch.  Rays are independent, so every view of a batch must come out
bit-identical to the same view rendered alone by gc_render_frame (FP32 and
FP64 compositing), with the batch's render counters the sum of the single
views' -- resident and out-of-core, eager and hipGraph replay."""
import numpy as np
import pytest

import gcvol as G
import oracle_lib as O
from scenes import Scene, tf_table
from test_gpu_parity import _compare_state

pytestmark = pytest.mark.gpu

SIZES = [(40, 30), (64, 64), (96, 80), (17, 23), (128, 72), (33, 65)]


def views_for(prec, tile=0, march=0, k=0, resume=0):
    out = []
    for i, (w, h) in enumerate(SIZES):
        cam = G.camera(15.0 + 7.0 * i, 25.0 * i + 11.0 * k, 2.6 + 0.15 * i)
        out.append(G.make_view(w, h, cam, alpha_stop=0.95 if i % 2 else 0.99, precision=prec, tile=tile, march=march,
                               resume=resume))
    return out


def one_pass(sc, views, d_tfs, sizes):
    """Each view in one launch over a fully loaded cache."""
    gc = sc.resident_gpu_cache()
    out = []
    for v, t, n in zip(views, d_tfs, sizes):
        f = G.Frame(v.width, v.height, v.precision)
        G.render(gc, v, t, n, f)
        out.append(f.pixels().reshape(-1).copy())
        f.close()
    gc.close()
    return out


def tfs(n):
    tables = [tf_table(seed=s) for s in (5, 6, 7)]
    d = [G.DeviceBuffer.from_array(t) for t in tables]
    return [d[i % 3] for i in range(n)], [len(tables[i % 3]) for i in range(n)]


def singles(cache, views, d_tfs, sizes):
    pix, st = [], np.zeros(4, np.uint64)
    for v, t, n in zip(views, d_tfs, sizes):
        f = G.Frame(v.width, v.height, v.precision)
        rep = G.render_frame(cache, v, t, n, f)
        assert rep.complete == 1
        pix.append(f.pixels().reshape(-1).copy())
        st += np.array([rep.stats.mapped_samples, rep.stats.brick_visits, rep.stats.pixels,
                        rep.stats.unmapped_rays], np.uint64)
    return pix, st


@pytest.mark.parametrize("prec,tile,march", [(32, 0, 0), (64, 2, 2), (32, 3, 1), (32, 4, 3), (64, 5, 4)])
def test_batch_equals_single_views_resident(gpu, prec, tile, march):
    sc = Scene((64, 64, 64), 8, 1, 1)
    gc = sc.gpu_cache()
    gc.load_all(sc.device_levels())
    views = views_for(prec, tile, march)
    d_tfs, sizes = tfs(len(views))
    print("batch resident prec", prec, "tile", tile, "march", march, flush=True)
    ref, ref_st = singles(gc, views, d_tfs, sizes)
    b = G.Batch(gc, 16, 100000, prec)
    rep = b.render(views, d_tfs, sizes)
    assert rep.complete == 1 and rep.launches == 1
    for i in range(len(views)):
        assert np.array_equal(b.pixels(i), ref[i]), "view %d differs" % i
    got = np.array([rep.stats.mapped_samples, rep.stats.brick_visits, rep.stats.pixels, rep.stats.unmapped_rays],
                   np.uint64)
    assert np.array_equal(got, ref_st)


@pytest.mark.parametrize("n_tables,entries", [(1, 256), (8, 256), (10, 256), (3, 300), (4, 64)])
def test_batch_transfer_function_tables(gpu, n_tables, entries):
    """Launches whose views share one table, use several LDS tables (up to 8
    of up to 256 entries) or fall back to reading the tables from HBM (more
    than 8 tables, or tables above 256 entries) render every view as it
    renders alone."""
    sc = Scene((64, 64, 64), 8, 1, 2)
    gc = sc.gpu_cache()
    gc.load_all(sc.device_levels())
    n_views = 12
    views = []
    for i in range(n_views):
        w, h = SIZES[i % len(SIZES)]
        cam = G.camera(10.0 + 5.0 * i, 31.0 * i, 2.5 + 0.1 * (i % 4))
        views.append(G.make_view(w, h, cam, alpha_stop=0.99, precision=32))
    tables = [tf_table(seed=40 + s, n=entries) for s in range(n_tables)]
    d = [G.DeviceBuffer.from_array(t) for t in tables]
    d_tfs = [d[i % n_tables] for i in range(n_views)]
    sizes = [entries] * n_views
    print("batch tf tables", n_tables, "entries", entries, flush=True)
    ref, ref_st = singles(gc, views, d_tfs, sizes)
    b = G.Batch(gc, 16, 100000, 32)
    rep = b.render(views, d_tfs, sizes)
    assert rep.complete == 1
    for i in range(n_views):
        assert np.array_equal(b.pixels(i), ref[i]), "view %d differs" % i


def _oracle_views(sc, views):
    tables = [tf_table(seed=s) for s in (5, 6, 7)]
    out = []
    for i, v in enumerate(views):
        cam = np.array(v.inv_mv[:], np.float32)
        out.append(O.make_view(sc.ocfg, v.width, v.height, cam, tables[i % 3], alpha_stop=v.alpha_stop))
    return out


@pytest.mark.parametrize("bytes_,graph,tile,handoff", [(1, False, 0, None), (2, True, 1, None), (2, False, 0, (1, 8, 0)),
                                                        (1, True, 0, (1, 16, 50)), (2, True, 0, (1, 8, 0)),
                                                        (2, False, 1, (0, 0, 0))])
def test_batch_out_of_core_lockstep(gpu, bytes_, graph, tile, handoff, parity_log):
    """Batched progressive launches with a pool smaller than the working set.
    After every launch, each view's pixels (<= 1 LSB), t buffer (exact) and
    FP64 colour state match the oracle rendering that view over the same cache
    state, the sample / visit / miss counters match exactly, the rays left
    waiting per view equal the oracle's misses, and every update leaves the
    LRU, page directory, usage, requests and pool bit-identical.  (The batch
    relaunches only the rays that stopped at a miss; the oracle re-runs every
    ray, whose finished rays neither sample nor touch the cache.)  handoff:
    the cache's hand-over setting (None: a new cache's; (1, lanes, 0): every
    ray that has sampled once is handed over to the group march at its next
    mapped brick; (0, 0, 0): off)."""
    sc = Scene((64, 64, 64), 8, 1, bytes_, cache_size=(9, 8, 8), max_requests=50)
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    if handoff is not None:
        G.set_ray_handoff(gc, *handoff)
    views = views_for(64, tile=tile)[:4]
    d_tfs, sizes = tfs(len(views))
    oviews = _oracle_views(sc, views)
    b = G.Batch(gc, 8, 8 * 128 * 80, 64)
    b.set_graph(graph)
    b.begin(views, d_tfs, sizes)
    otb = [np.zeros(v.width * v.height, np.float32) for v in views]
    ocs = [np.zeros(v.width * v.height * 4) for v in views]
    prev = np.zeros(4, np.uint64)
    launches = 0
    errs = []
    for it in range(80):
        print("lockstep launch", it, flush=True)
        b.launch()
        launches += 1
        st = b.stats()
        now = np.array([st.mapped_samples, st.brick_visits, st.pixels, st.unmapped_rays], np.uint64)
        delta, prev = now - prev, now
        pend = b.pending()
        ototal = np.zeros(4, np.uint64)
        for i, v in enumerate(views):
            opx = np.zeros(v.width * v.height, np.uint32)
            ost = oc.render(oviews[i], otb[i], ocs[i], opx, precision=64, nthreads=8)
            ototal += ost
            err = np.abs(G.unpack_rgba(b.pixels(i)) - G.unpack_rgba(opx))
            errs.append(err.reshape(-1))
            assert err.max() <= 1, (it, i)
            tb, cs = b.state(i)
            assert np.array_equal(tb, otb[i]), (it, i)
            assert np.allclose(cs, ocs[i], rtol=1e-12, atol=1e-15), (it, i)
            assert pend[i] == ost[3], (it, i)
        assert delta[[0, 1, 3]].tolist() == ototal[[0, 1, 3]].tolist(), it
        if it == 0:
            assert delta[2] == ototal[2]
        gd, od = gc.update(), oc.update()
        assert gd == od
        _compare_state(gc, oc, pool=(it % 4 == 0))
        if gd:
            break
    assert gd and launches > 3
    parity_log(64, np.concatenate(errs), launches=launches)


def test_batch_graph_replay_matches_eager(gpu):
    sc = Scene((64, 64, 64), 8, 1, 2)
    gc = sc.gpu_cache()
    gc.load_all(sc.device_levels())
    d_tfs, sizes = tfs(len(SIZES))
    eager = G.Batch(gc, 16, 100000, 32)
    graph = G.Batch(gc, 16, 100000, 32)
    graph.set_graph(True)
    for k in range(4):
        views = views_for(32, k=k)       # new cameras, same view count: the captured graph is reused
        print("graph replay", k, flush=True)
        re = eager.render(views, d_tfs, sizes)
        rg = graph.render(views, d_tfs, sizes)
        assert rg.complete == 1 and rg.stats.mapped_samples == re.stats.mapped_samples
        for i in range(len(views)):
            assert np.array_equal(graph.pixels(i), eager.pixels(i)), (k, i)
    # a different view count re-captures
    views = views_for(32)[:3]
    re = eager.render(views, d_tfs[:3], sizes[:3])
    rg = graph.render(views, d_tfs[:3], sizes[:3])
    for i in range(3):
        assert np.array_equal(graph.pixels(i), eager.pixels(i))


def test_upload_batch_limit(gpu):
    sc = Scene((64, 64, 64), 8, 1, 1, max_requests=64)
    gc = sc.gpu_cache()
    with pytest.raises(G.GcError):
        G.set_max_requests(gc, 65)
    G.set_max_requests(gc, 7)
    view = G.make_view(64, 64, G.camera(20.0, 30.0, 3.0), precision=32)
    tf = tf_table()
    d_tf = G.DeviceBuffer.from_array(tf)
    f = G.Frame(64, 64, 32)
    rep = G.render_frame(gc, view, d_tf, len(tf), f)
    assert rep.complete == 1
    asked, _ = gc.asked()
    assert 0 < len(asked) <= 7


def test_starting_only_launch_keeps_waiting_rays(gpu):
    """A launch that traces only the starting views leaves the rays waiting
    for bricks in their list (their per-view counts unchanged) and appends
    the new view's stopped rays; the views then complete as usual.  With
    exact continuation each view equals its one-pass render bit for bit."""
    sc = Scene((64, 64, 64), 8, 1, 1, max_requests=50)
    gc = sc.gpu_cache()
    views = views_for(32, resume=1)[:2]
    d_tfs, sizes = tfs(2)
    b = G.Batch(gc, 2, 2 * 128 * 80, 32)
    b.set_view(0, views[0], d_tfs[0], sizes[0])
    b.launch()
    p0 = b.pending()
    assert p0[0] > 0 and p0[1] == 0
    b.set_view(1, views[1], d_tfs[1], sizes[1])
    b.launch(starting_only=True)
    p1 = b.pending()
    assert p1[0] == p0[0] and p1[1] > 0
    for it in range(200):
        gc.update()
        b.launch()
        if not b.pending().any():
            break
    assert not b.pending().any()
    ref = one_pass(sc, views, d_tfs, sizes)
    for i in range(2):
        assert np.array_equal(b.pixels(i), ref[i]), i


@pytest.mark.parametrize("prec,graph", [(32, False), (64, True)])
def test_batch_exact_continuation_lockstep(gpu, prec, graph, parity_log):
    """Progressive batched launches with exact continuation (resume = 1) and
    a pool smaller than the working set, in lockstep with the oracle's
    exact-continuation mode: after every launch the pixels are within 1/255
    of the FP64 oracle, the t buffers and counters are equal, and every
    update leaves the cache state bit-identical; at the end every view equals
    its one-pass render bit for bit."""
    sc = Scene((64, 64, 64), 8, 1, 2, cache_size=(9, 8, 8), max_requests=50)
    gc, oc = sc.gpu_cache(), sc.oracle_cache()
    views = views_for(prec, resume=1)[:4]
    d_tfs, sizes = tfs(len(views))
    oviews = _oracle_views(sc, views)
    b = G.Batch(gc, 8, 8 * 128 * 80, prec)
    b.set_graph(graph)
    b.begin(views, d_tfs, sizes)
    otb = [np.zeros(v.width * v.height, np.float32) for v in views]
    ocs = [np.zeros(v.width * v.height * 4) for v in views]
    orp = [np.zeros(v.width * v.height * 4, np.float32) for v in views]
    prev = np.zeros(4, np.uint64)
    errs = []
    for it in range(80):
        print("exact continuation lockstep launch", it, flush=True)
        b.launch()
        st = b.stats()
        now = np.array([st.mapped_samples, st.brick_visits, st.pixels, st.unmapped_rays], np.uint64)
        delta, prev = now - prev, now
        pend = b.pending()
        ototal = np.zeros(4, np.uint64)
        for i, v in enumerate(views):
            opx = np.zeros(v.width * v.height, np.uint32)
            ost = oc.render(oviews[i], otb[i], ocs[i], opx, precision=64, nthreads=8, rpos=orp[i])
            ototal += ost
            err = np.abs(G.unpack_rgba(b.pixels(i)) - G.unpack_rgba(opx))
            errs.append(err.reshape(-1))
            assert err.max() <= 1, (it, i)
            tb, cs = b.state(i)
            assert np.array_equal(tb, otb[i]), (it, i)
            assert pend[i] == ost[3], (it, i)
        assert delta[[0, 1, 3]].tolist() == ototal[[0, 1, 3]].tolist(), it
        gd, od = gc.update(), oc.update()
        assert gd == od
        _compare_state(gc, oc, pool=(it % 4 == 0))
        if gd:
            break
    assert gd and it > 3
    ref = one_pass(sc, views, d_tfs, sizes)
    for i in range(len(views)):
        assert np.array_equal(b.pixels(i), ref[i]), i
    parity_log(prec, np.concatenate(errs), launches=it + 1)
