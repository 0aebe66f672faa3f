"""Hand-over of long rays in the frame loop's lane-per-ray launches
(gc_cache_set_ray_handoff): a ray that has consumed the budget of samples in
a launch stops before the first sample of its next mapped brick, and a group
march continues it from the stored state within the same launch, before the
cache update.  The same out-of-core frame sequence rendered with and without
it, on two fresh caches, must give bit-identical pixels, t buffers, colour
state, render counters and cache state (page directory, LRU, usage and
request buffers) after every frame.  The launch log must show rays handed
over, and its records must still sum to the frame report."""
import numpy as np
import pytest

import gcvol as G
from scenes import Scene, orbit, tf_table

pytestmark = pytest.mark.gpu


# (voxel bytes, brick size, precision, budget, lanes per ray of the group
# march, percentage of exited waves before rays are handed over, resumed
# launches: "lane" = always the lane march, "default" = the group march for
# short lists)
CASES = [(1, 8, 32, 4, 16, 0, "lane"), (2, 8, 64, 16, 64, 0, "lane"), (2, 8, 32, 1, 4, 0, "lane"),
         (2, 32, 32, 8, 16, 0, "lane"), (1, 32, 32, 24, 8, 0, "lane"), (2, 32, 32, 8, 16, 0, "default"),
         (2, 8, 32, 1, 16, 50, "lane"), (1, 32, 32, 1, 8, 90, "lane"), (2, 32, 64, 2, 16, 25, "default")]


@pytest.mark.parametrize("bytes_,bs,prec,budget,lanes,exited,resume", CASES)
def test_handoff_matches_lane_march(gpu, bytes_, bs, prec, budget, lanes, exited, resume):
    if bs == 8:
        sc = Scene((96, 80, 64), 8, 1, bytes_, cache_size=(9, 9, 8), max_requests=64)
        W, H = 80, 64
    else:
        sc = Scene((160, 128, 96), 32, 1, bytes_, cache_size=(5, 4, 4), max_requests=64)
        W, H = 96, 80
    caches = [sc.gpu_cache(), sc.gpu_cache()]
    assert G.get_ray_handoff(caches[0]) == (32, 8, 75)   # a new cache's setting
    if resume == "lane":
        for c in caches:
            G.set_ray_list_max(c, 0)
    G.set_ray_handoff(caches[0], 0)             # no hand-over (a new cache has it on)
    G.set_ray_handoff(caches[1], budget, lanes, exited)
    log = G.LaunchLog(caches[1], capacity=512)
    tf = tf_table(seed=23)
    d_tf = G.DeviceBuffer.from_array(tf)
    handed = 0
    for k in range(4):
        out = []
        for c in caches:
            f = G.Frame(W, H, prec)
            view = G.make_view(W, H, orbit(2 * k, distance=2.2 if bs == 8 else 2.0 + 0.3 * k),
                               alpha_stop=0.95 if k % 2 else 0.99, precision=prec)
            r = G.render_frame(c, view, d_tf, len(tf), f)
            tb, cs = f.state()
            out.append((f.pixels().copy(), tb.copy(), cs.copy(), r.as_dict()))
            f.close()
        (p0, t0, c0, r0), (p1, t1, c1, r1) = out
        recs = log.last_frame()
        n = sum(r["handed_over"] for r in recs)
        handed += n
        print("handoff frame", k, "launches", r0["launches"], "samples", r0["mapped_samples"], "handed over", n,
              flush=True)
        assert r0["complete"] and r1["complete"] and r0["launches"] > 1
        for key in ("launches", "mapped_samples", "brick_visits", "pixels", "unmapped_rays", "bricks_loaded"):
            assert r0[key] == r1[key], (k, key, r0[key], r1[key])
        assert np.array_equal(p0, p1), "pixels differ at frame %d" % k
        assert np.array_equal(t0, t1) and np.array_equal(c0.view(np.uint8), c1.view(np.uint8)), k
        a, b = caches
        assert np.array_equal(a.mrpd(), b.mrpd()) and np.array_equal(a.lru(), b.lru()), k
        assert np.array_equal(a.usage(), b.usage()) and np.array_equal(a.request(), b.request()), k
        # the log of the frame with hand-over still sums to its report
        assert len(recs) == r1["launches"]
        for key in ("mapped_samples", "brick_visits", "unmapped_rays"):
            assert sum(r[key] for r in recs) == r1[key], key
        assert sum(r["rays"] for r in recs) == r1["pixels"]
        for x, y in zip(recs, recs[1:]):
            assert y["list_rays"] == x["unmapped_rays"] == y["rays"]
        assert all(r["handed_over"] == 0 for r in recs if r["kernel"] != "lane")
    assert handed > 0
    log.close()
    for c in caches:
        c.close()
