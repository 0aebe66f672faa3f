"""The per-launch log of the frame loop (gc_cache_set_launch_log): one record
per launch, whose counters sum to the frame report's and whose kernels match
its split between the tile / lane march and the group march."""
import numpy as np
import pytest

import gcvol as G
from scenes import Scene, orbit, tf_table

pytestmark = pytest.mark.gpu


def test_launch_log_sums_to_frame_report(gpu):
    sc = Scene((96, 80, 64), 8, 1, 2, cache_size=(9, 9, 8), max_requests=64)
    gc = sc.gpu_cache()
    log = G.LaunchLog(gc, capacity=256)
    tf = tf_table(seed=7)
    d_tf = G.DeviceBuffer.from_array(tf)
    for k in range(3):
        f = G.Frame(80, 64, 32)
        rep = G.render_frame(gc, G.make_view(80, 64, orbit(4 * k, distance=2.2), alpha_stop=0.99), d_tf, len(tf), f)
        d = rep.as_dict()
        recs = log.last_frame()
        print("frame", k, "launches", d["launches"], [(r["kernel"], r["list_rays"], r["rays"]) for r in recs][:6],
              flush=True)
        assert len(recs) == d["launches"] > 1
        assert [r["launch"] for r in recs] == list(range(d["launches"]))
        assert recs[0]["list_rays"] == 0 and recs[0]["kernel"] == "lane"
        for key in ("mapped_samples", "brick_visits", "unmapped_rays"):
            assert sum(r[key] for r in recs) == d[key], key
        assert sum(r["rays"] for r in recs) == d["pixels"]
        assert sum(r["bricks_loaded"] for r in recs) == d["bricks_loaded"]
        assert sum(r["kernel"] == "lane" for r in recs) == d["lane_launches"]
        assert all(r["group"] in (4, 8, 16, 64) for r in recs if r["kernel"] == "group")
        # every resumed launch starts the rays the previous one stopped
        for a, b in zip(recs, recs[1:]):
            assert b["list_rays"] == a["unmapped_rays"] == b["rays"]
        assert abs(sum(r["kernel_ms"] for r in recs) - d["kernel_ms"]) < 1e-3 * max(1.0, d["kernel_ms"])
        f.close()
    log.close()
    gc.close()
