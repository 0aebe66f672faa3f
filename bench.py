#!/usr/bin/env python3
"""Out-of-core DVR benchmark on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on):
a 2048^3 uint16 volume synthesised brick by brick on the host from a seeded
hash-noise + 64 Gaussian-blob function (standing in for the brick store),
32^3 bricks with 1 covering voxel, 7 levels, a 4 GiB device brick pool (LRU
eviction exercised), 1024x1024 views along a 120-frame fly-through, step 0.5
voxel of the selected level, 256-entry seeded transfer function, early
termination at alpha 0.99.

A step is one frame: the plain single-view renderer's progressive loop
(launch, update the cache, load the requested bricks, relaunch) until the
view is complete.  `value` is ray samples per second over the timed frames,
wall clock, including the on-demand brick synthesis and upload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-out-of-core-volume-data_amd", "python"))

METRIC = "ray samples/s + frames/s, p50/p99 frame latency; batched vs plain renderer"
HBM_PEAK = 8.0e12

DIMS = (2048, 2048, 2048)
BRICK = 32
COVER = 1
BYTES = 2
POOL_BYTES = 4 << 30
WIDTH = HEIGHT = 1024
PATH_FRAMES = 120
ALPHA_STOP = 0.99
SEED = 20240611


def fly_through(k):
    """Frame k of the 120-frame camera path: orbit 3 degrees per frame while
    closing in from distance 3.2 to 1.6, pitch oscillating 10-30 degrees."""
    import gcvol as G

    s = k / (PATH_FRAMES - 1)
    rx = 20.0 + 10.0 * math.sin(2.0 * math.pi * s)
    ry = 3.0 * k
    dist = 3.2 - 1.6 * s
    return G.camera(rx, ry, dist)


def algorithmic_bytes(stats, bytes_per_voxel):
    """B = S*8*sizeof(T) + V*(16+4) + P*44 (box footprint per mapped sample,
    page entry + usage stamp per brick visit, RGBA8 + t read/write + colour
    state read/write per pixel)."""
    return (stats["mapped_samples"] * 8 * bytes_per_voxel + stats["brick_visits"] * 20 + stats["pixels"] * 44)


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return float("nan")
    k = (len(xs) - 1) * q / 100.0
    lo, hi = int(math.floor(k)), int(math.ceil(k))
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def cpu_baseline(G, cache, vc, cam, tf, seconds=12.0):
    """The oracle's FP64 host ray caster (test infrastructure) on the host
    cores, over the device cache state of the last timed frame: the same
    page directory and pool, the same 1024x1024 view.  Rows are rendered in
    bands until about `seconds` of CPU work; samples/s is reported."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    cores = os.cpu_count() or 1
    g = cache.g
    ocfg = O.make_cfg(vc.initial(), vc.real(), BRICK, COVER, BYTES,
                      [int(g.cache_size[a]) for a in range(3)])
    oc = O.Cache(ocfg, [np.zeros(1, np.uint8)] * vc.nb_levels)
    oc.mrpd()[:] = cache.mrpd()
    # device pool straight into the oracle's 3-D pool
    G.check(G.lib().gc_cache_read_pool(cache.h, O.lib().or_cache_pool(oc.h)))
    band = 64
    rows_done = 0
    samples = 0
    t0 = time.perf_counter()
    while rows_done < HEIGHT and time.perf_counter() - t0 < seconds:
        h = min(band, HEIGHT - rows_done)
        # render rows [rows_done, rows_done+h) through a sub-view with the
        # same camera: the oracle's rays depend on (x, y, W, H) only, so the
        # band is rendered as part of the full frame by offsetting pixels.
        view = O.make_view(ocfg, WIDTH, HEIGHT, cam, tf, alpha_stop=ALPHA_STOP)
        tb = np.zeros(WIDTH * HEIGHT, np.float32)
        cs = np.zeros(WIDTH * HEIGHT * 4)
        px = np.zeros(WIDTH * HEIGHT, np.uint32)
        st = render_rows(O, oc, view, tb, cs, px, rows_done, h, cores)
        samples += int(st[0])
        rows_done += h
    dt = time.perf_counter() - t0
    oc.close()
    return {"value": samples / dt, "unit": "samples/s", "cores": cores, "kind": "port",
            "sample": "FP64 oracle ray caster, rows 0-%d of the last timed 1024x1024 view over the device "
                      "cache state (%.1f s, %d samples)" % (rows_done, dt, samples)}


def render_rows(O, oc, view, tb, cs, px, y0, h, threads):
    import ctypes as C

    st = np.zeros(4, np.uint64)
    O.lib().or_render_dvr_rows.argtypes = [C.c_void_p, C.POINTER(O.OrView), C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_void_p]
    O.lib().or_render_dvr_rows(oc.h, C.byref(view), O.ptr(tb), O.ptr(cs), O.ptr(px), 64, threads, y0, h,
                               O.ptr(st))
    return st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--max-requests", type=int, default=4096)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--precision", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a kernel of known traffic "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import gcvol as G

    if G.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)

    if args.calib_copy_mib > 0:
        nbytes = args.calib_copy_mib << 20
        a, b = G.DeviceBuffer(nbytes), G.DeviceBuffer(nbytes)
        G.check(G.lib().gc_memset(a.p, 1, nbytes))
        G.check(G.lib().gc_memcpy_d2d(b.p, a.p, nbytes))
        G.synchronize()
        a.free()
        b.free()

    vc = G.volume_cfg(DIMS, BRICK, COVER, BYTES)
    params = G.synth_params(SEED, 64, sigma=(40.0, 200.0), amp=(15000.0, 65535.0), noise=2048, threads=0)
    src = G.Source.synthetic(vc, params)
    brick_bytes = BRICK ** 3 * BYTES
    cs = G.pool_shape(POOL_BYTES, brick_bytes)
    cc = G.cache_cfg(cs, BRICK, max_requests=args.max_requests, sync=False)
    cache = G.Cache(vc, cc, src)
    tf = G.tf_samples(G.tf_seeded(SEED, 6), 256)
    d_tf = G.DeviceBuffer.from_array(tf)
    frame = G.Frame(WIDTH, HEIGHT, args.precision)

    nframes = args.warmup + args.steps
    lat, reports = [], []
    cams = []
    for k in range(nframes):
        cam = fly_through(k % PATH_FRAMES)
        cams.append(cam)
        view = G.make_view(WIDTH, HEIGHT, cam, alpha_stop=ALPHA_STOP, precision=args.precision, tile=args.tile)
        G.synchronize()
        t0 = time.perf_counter()
        frame.reset()
        rep = G.render_frame(cache, view, d_tf, len(tf), frame)
        dt = time.perf_counter() - t0
        d = rep.as_dict()
        d["wall_ms"] = dt * 1e3
        print("frame %3d %s wall %.1f ms launches %d bricks %d samples %.3g kernel %.1f ms" % (
            k, "timed" if k >= args.warmup else "warmup", dt * 1e3, d["launches"], d["bricks_loaded"],
            d["mapped_samples"], d["kernel_ms"]), file=sys.stderr, flush=True)
        if k >= args.warmup:
            lat.append(dt * 1e3)
            reports.append(d)
    total_s = sum(lat) / 1e3
    S = sum(r["mapped_samples"] for r in reports)
    kms = sum(r["kernel_ms"] for r in reports)
    launches = sum(r["launches"] for r in reports)
    tot = {k: sum(r[k] for r in reports) for k in ("mapped_samples", "brick_visits", "pixels")}
    B = algorithmic_bytes(tot, BYTES)
    achieved = B / (kms / 1e3) if kms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": S / total_s if total_s > 0 else 0.0,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sum(lat) / len(lat) if lat else 0.0,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic",
        "config": {
            "workload": "out-of-core single view: 2048^3 u16 synthesised on demand (seeded hash noise + 64 blobs), "
                        "32^3 bricks, covering 1, 7 levels, 4 GiB pool (%dx%dx%d slots), 1024x1024, 120-frame "
                        "fly-through, step 0.5 voxel, alpha stop 0.99, %d requests/update" % (cs + (args.max_requests,)),
            "precision": "fp%d compositing" % args.precision,
            "tile": ["8x8", "16x8", "32x4", "64x1"][args.tile],
        },
        "frames_per_s": len(lat) / total_s if total_s > 0 else 0.0,
        "p50_ms": percentile(lat, 50),
        "p99_ms": percentile(lat, 99),
        "kernel_samples_per_s": S / (kms / 1e3) if kms > 0 else 0.0,
        "launches": launches,
        "bricks_loaded": sum(r["bricks_loaded"] for r in reports),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": None,
            "kernel": "k_dvr<2,%s,%s>" % ("double" if args.precision == 64 else "float", ["8,8", "16,4", "32,2", "64,1"][args.tile]),
            "avg_launch_ms": kms / launches if launches else 0.0,
            "algorithmic_bytes_per_launch": B / launches if launches else 0.0,
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(G, cache, vc, cams[-1], tf, seconds=args.cpu_seconds)
    cache.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
