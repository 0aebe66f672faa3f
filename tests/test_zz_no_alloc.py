"""gc_render_views_noalloc (one launch over caller buffers, scratch in a
per-device static array) against the entry points it varies,
gc_batch_begin + gc_batch_launch: two calls in a row, the second with fewer
and smaller views, on a resident cache and on an out-of-core cache whose
rays stop at missing bricks.  Pixels, t buffers and colour states must be
bit-identical, and so must the cache's usage and request buffers after each
launch (the two paths run on two identical caches)."""
import numpy as np
import pytest

import gcvol as G
from scenes import Scene, tf_table

pytestmark = pytest.mark.gpu

SIZES = [(96, 80), (40, 30), (64, 64), (17, 23), (128, 72)]


def _views(sizes, prec, k):
    out = []
    for i, (w, h) in enumerate(sizes):
        cam = G.camera(12.0 + 9.0 * i, 31.0 * i + 17.0 * k, 2.5 + 0.2 * i)
        out.append(G.make_view(w, h, cam, alpha_stop=0.99 if i % 2 else 0.95, precision=prec))
    return out


@pytest.mark.parametrize("prec,resident", [(32, True), (64, True), (32, False)])
def test_noalloc_matches_batch_launch(gpu, prec, resident):
    sc = Scene((64, 64, 64), 8, 1, 2, max_requests=50, cache_size=None if resident else (6, 6, 5))
    ref_c, na_c = sc.gpu_cache(), sc.gpu_cache()
    if resident:
        ref_c.load_all(sc.device_levels())
        na_c.load_all(sc.device_levels())
    tables = [tf_table(seed=s) for s in (5, 9)]
    d_tab = [G.DeviceBuffer.from_array(t) for t in tables]
    cwords = 8 if prec == 64 else 4
    b = G.Batch(ref_c, 8, 8 * 128 * 80, prec)
    for call, sizes in enumerate((SIZES, SIZES[1:3])):
        views = _views(sizes, prec, call)
        d_tfs = [d_tab[i % 2] for i in range(len(views))]
        tf_sizes = [len(tables[i % 2]) for i in range(len(views))]
        print("noalloc call", call, "views", len(views), "precision", prec, "resident", resident, flush=True)
        b.begin(views, d_tfs, tf_sizes)
        b.launch()
        px = [G.DeviceBuffer(w * h * 4) for w, h in sizes]
        tb = [G.DeviceBuffer(w * h * 4) for w, h in sizes]
        cs = [G.DeviceBuffer(w * h * 4 * cwords) for w, h in sizes]
        G.render_views_noalloc(na_c, views, d_tfs, tf_sizes, px, tb, cs)
        G.synchronize()
        for i, (w, h) in enumerate(sizes):
            assert np.array_equal(px[i].to_array(np.uint32, (w * h,)), b.pixels(i)), (call, i)
            rt, rc = b.state(i)
            assert np.array_equal(tb[i].to_array(np.float32, (w * h,)), rt), (call, i)
            got_c = cs[i].to_array(np.float64 if prec == 64 else np.float32, (w * h * 4,))
            assert np.array_equal(got_c.view(np.uint8), rc.view(np.uint8)), (call, i)
        assert np.array_equal(ref_c.usage(), na_c.usage()), call
        assert np.array_equal(ref_c.request(), na_c.request()), call
        if not resident:
            assert b.pending()[:len(views)].sum() > 0      # rays stopped at missing bricks
            ref_c.update()
            na_c.update()
    with pytest.raises(G.GcError):
        many = _views([(8, 8)] * 65, prec, 0)
        bufs = [G.DeviceBuffer(8 * 8 * 4 * cwords) for _ in many]
        G.render_views_noalloc(na_c, many, [d_tab[0]] * 65, [len(tables[0])] * 65, bufs, bufs, bufs)
