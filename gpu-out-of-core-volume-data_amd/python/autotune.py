"""Autotuner for the ray caster's launch shape and the cache's upload batch.

Knobs (all runtime parameters of the C ABI, no rebuild):
  * tile   -- shape of a 64-ray packet: 0 8x8, 1 16x4, 2 32x2, 3 64x1;
  * march  -- samples per lane between refill checks (1..8);
  * upload batch -- bricks requested per cache update (gc_cache_set_max_requests).

tune_launch() times every (tile, march) pair on a warm batch -- each timing
is the kernel time of a launch that renders the whole batch, the best of
`reps` -- and returns the fastest.  tune_upload() replays the same cold frames
with each upload batch on a fresh cache and keeps the one with the lowest mean
frame time.  Results are plain dicts so callers can persist them.
"""
import time

import numpy as np

import gcvol as G


def tune_launch(cache, views, d_tfs, tf_sizes, tiles=(0, 1, 2, 3), marches=(1, 2, 3, 4), reps=3,
                precision=32):
    """Best (tile, march) for rendering `views` together over `cache`.
    The cache should already hold the views' bricks (render once first)."""
    npix = sum(v.width * v.height for v in views)
    b = G.Batch(cache, max(len(views), 1), npix, precision)
    table = []
    try:
        b.render(views, d_tfs, tf_sizes)          # warm: every brick the views touch is resident
        for tile in tiles:
            for march in marches:
                for v in views:
                    v.tile, v.march = tile, march
                best = float("inf")
                samples = 0
                for _ in range(reps):
                    rep = b.render(views, d_tfs, tf_sizes)
                    if rep.launches == 1:
                        best = min(best, rep.kernel_ms)
                    samples = rep.stats.mapped_samples
                table.append({"tile": tile, "march": march, "kernel_ms": best,
                              "samples_per_s": samples / (best / 1e3) if best < float("inf") else 0.0})
    finally:
        b.close()
    table.sort(key=lambda r: r["kernel_ms"])
    return {"tile": table[0]["tile"], "march": table[0]["march"], "table": table}


def tune_upload(make_cache, frames, candidates=(1024, 2048, 4096, 8192)):
    """Best upload batch (bricks per update) for a cold sequence of frames.
    make_cache() returns a fresh cache created with max_requests >=
    max(candidates); frames is a list of (view, d_tf, tf_size) rendered in
    order with gc_render_frame."""
    table = []
    for n in candidates:
        cache = make_cache()
        try:
            G.set_max_requests(cache, n)
            times = []
            for view, d_tf, tf_size in frames:
                f = G.Frame(view.width, view.height, view.precision)
                t0 = time.perf_counter()
                G.render_frame(cache, view, d_tf, tf_size, f)
                times.append((time.perf_counter() - t0) * 1e3)
                f.close()
        finally:
            cache.close()
        table.append({"max_requests": n, "mean_frame_ms": float(np.mean(times)), "max_frame_ms": float(max(times))})
    table.sort(key=lambda r: r["mean_frame_ms"])
    return {"max_requests": table[0]["max_requests"], "table": table}
