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
view is complete.  The K timed frames are spread evenly over the 120-frame
path (frame j of K is path frame j*120/K), so they cover the far start and
the close-up end; the W warm-up frames are the path's first W frames.
`value` is ray samples per second over the timed frames, wall clock,
including the on-demand brick synthesis and upload.

Beside the headline line, the same JSON carries: the plain-method
configuration (256^3 u8 blobs, resident, 36-frame 512^2 orbit) on the GPU
and on the host cores (bricker + integer mip + FP64 caster), the batched vs
plain renderer on 64 uneven views, and a bounded serving run at 50, 200 and
800 requests/s (batched serving loop vs one request at a time).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-traffic]

roofline.traffic is the ray caster's HBM bytes per launch, measured before
the timed run by two short child runs of this benchmark under
`rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE`, each scaled by what the
counter reads for a device-to-device copy of known size in the same run.
"""
import argparse
import json
import math
import os
import shutil
import sqlite3
import subprocess
import sys
import tempfile
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


def path_frame(k, warmup, steps):
    """Path frame of run frame k: warm-up frames are the path's first
    frames, timed frame j of `steps` is path frame j * 120 / steps."""
    if k < warmup:
        return k % PATH_FRAMES
    return ((k - warmup) * PATH_FRAMES // max(1, steps)) % PATH_FRAMES


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


def cpu_baseline(G, cache, vc, cam, tf, seconds=12.0, device_pixels=None):
    """The oracle's FP64 host ray caster (test infrastructure) on the host
    cores, over the device cache state right after the first timed frame
    (path frame 0, whose bricks all fit in the pool, so the state holds every
    brick the view needs): the same page directory and pool, the same
    1024x1024 view.  Rows are rendered in bands until about `seconds` of CPU
    work; samples/s is reported, and the rendered rows' largest RGBA
    difference from the device frame."""
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
    samples = unmapped = 0
    maxdiff = 0
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
        unmapped += int(st[3])
        if device_pixels is not None:
            a = px[rows_done * WIDTH:(rows_done + h) * WIDTH]
            b = device_pixels[rows_done * WIDTH:(rows_done + h) * WIDTH]
            maxdiff = max(maxdiff, int(np.abs(G.unpack_rgba(a) - G.unpack_rgba(b)).max()))
        rows_done += h
    dt = time.perf_counter() - t0
    oc.close()
    return {"value": samples / dt, "unit": "samples/s", "cores": cores, "kind": "port",
            "sample": "FP64 oracle ray caster, rows 0-%d of the first timed 1024x1024 view (path frame 0) over the "
                      "device cache state after it (%.1f s, %d samples, %d rays stopped at a missing brick)"
                      % (rows_done, dt, samples, unmapped),
            "max_rgba_diff_vs_device_frame": maxdiff if device_pixels is not None else None}


def render_rows(O, oc, view, tb, cs, px, y0, h, threads):
    import ctypes as C

    st = np.zeros(4, np.uint64)
    O.lib().or_render_dvr_rows.argtypes = [C.c_void_p, C.POINTER(O.OrView), C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_void_p]
    O.lib().or_render_dvr_rows(oc.h, C.byref(view), O.ptr(tb), O.ptr(cs), O.ptr(px), 64, threads, y0, h,
                               O.ptr(st))
    return st


CALIB_MIB = 256
TILE_NAMES = ["8x8", "16x4", "32x2", "64x1", "16x8", "32x4"]

# plain-method configuration: 256^3 u8, 12 blobs + noise, 16^3 bricks, resident
C1_DIMS = (256, 256, 256)
C1_BRICK = 16
C1_SIZE = 512
C1_FRAMES = 36


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    return O


def cpu_mip_and_bricks(O, vol, nlev, bs, cov, cores):
    """The CPU preprocessor on `cores` threads: integer 2x2x2 mips level after
    level (z-slabs of a level in parallel) and the bricker's cut of every level
    (brick ranges in parallel).  Returns (levels, mip seconds, brick seconds,
    bricks)."""
    from concurrent.futures import ThreadPoolExecutor

    levels = [vol]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        for _ in range(1, nlev):
            src = levels[-1]
            Z, Y, X = src.shape
            oz = Z // 2
            out = np.zeros((oz, Y // 2, X // 2), src.dtype)
            step = max(1, (oz + cores - 1) // cores)

            def slab(z0, src=src, out=out):
                z1 = min(oz, z0 + step)
                out[z0:z1] = O.mip_reduce(src[2 * z0:2 * z1])

            list(ex.map(slab, range(0, oz, step)))
            levels.append(out)
        t_mip = time.perf_counter() - t0
        t0 = time.perf_counter()
        inner = bs - 2 * cov
        bricks = 0
        jobs = []
        for lv in levels:
            nb = [(d + inner - 1) // inner for d in (lv.shape[2], lv.shape[1], lv.shape[0])]
            n = nb[0] * nb[1] * nb[2]
            bricks += n
            step = max(1, (n + cores - 1) // cores)
            jobs += [(lv, b0, min(n, b0 + step)) for b0 in range(0, n, step)]
        list(ex.map(lambda j: O.brick_level(j[0], bs, cov, j[1], j[2]), jobs))
        t_brick = time.perf_counter() - t0
    return levels, t_mip, t_brick, bricks


def plain_config_workload(G, cpu_frames=2):
    """The plain-method configuration: a 256^3 u8 volume of 12 seeded blobs
    (sigma 8-40, amplitude 60-255) plus noise +-8, 16^3 bricks with covering
    1, 5 levels, fully resident, 512x512 views on the 36-frame orbit (10
    degrees per frame, distance 3), alpha stop 0.99.  GPU: mips and brick
    cuts on the device, the resident load, 36 one-launch frames (FP32
    compositing).  CPU (the oracle, test infrastructure, on all host cores):
    the integer mips, the bricker's cut of every level and the FP64 caster on
    `cpu_frames` of the frames, whose pixels are also compared with the GPU's."""
    O = _oracle()
    cores = os.cpu_count() or 1
    params = G.synth_params(SEED + 1, 12, sigma=(8.0, 40.0), amp=(60.0, 255.0), noise=8, threads=0)
    vol = G.synth_volume(params, C1_DIMS, 1)
    vc = G.volume_cfg(C1_DIMS, C1_BRICK, 1, 1)
    nl = vc.nb_levels
    # GPU preprocessor: mips, then every level's bricks
    d_levels, dims = [G.DeviceBuffer.from_array(vol)], [C1_DIMS]
    G.synchronize()
    t0 = time.perf_counter()
    for _ in range(1, nl):
        nd = tuple(d // 2 for d in dims[-1])
        d_out = G.DeviceBuffer(max(16, nd[0] * nd[1] * nd[2]))
        G.mip_reduce(d_levels[-1], dims[-1], 1, d_out)
        d_levels.append(d_out)
        dims.append(nd)
    G.synchronize()
    gpu_mip_ms = (time.perf_counter() - t0) * 1e3
    inner = C1_BRICK - 2
    lists, outs, counts = [], [], []
    for d in dims:
        nb = [(x + inner - 1) // inner for x in d]
        br = np.array([[x, y, z] for z in range(nb[2]) for y in range(nb[1]) for x in range(nb[0])], np.uint32)
        lists.append(G.DeviceBuffer.from_array(br))
        outs.append(G.DeviceBuffer(len(br) * C1_BRICK ** 3))
        counts.append(len(br))
    total = sum(counts)
    G.synchronize()
    t0 = time.perf_counter()
    for l, d in enumerate(dims):
        G.check(G.lib().gc_brick_cut(d_levels[l].p, G.ptr(np.asarray(d, np.uint32)), 1, C1_BRICK, 1, lists[l].p,
                                     counts[l], outs[l].p, None))
    G.synchronize()
    gpu_brick_ms = (time.perf_counter() - t0) * 1e3
    for b in lists + outs:
        b.free()
    host_levels = [d_levels[l].to_array(np.uint8, (dims[l][2], dims[l][1], dims[l][0])) for l in range(nl)]
    cs = G.pool_shape(total * C1_BRICK ** 3, C1_BRICK ** 3)
    cache = G.Cache(vc, G.cache_cfg(cs, C1_BRICK, max_requests=64, sync=True), G.Source.pyramid(vc, host_levels))
    G.synchronize()
    t0 = time.perf_counter()
    cache.load_all(d_levels)
    G.synchronize()
    load_ms = (time.perf_counter() - t0) * 1e3
    tf = G.tf_samples(G.tf_seeded(SEED + 1, 6), 256)
    d_tf = G.DeviceBuffer.from_array(tf)
    frame = G.Frame(C1_SIZE, C1_SIZE, 32)
    d_st = G.DeviceBuffer.from_array(np.zeros(4, np.uint64))
    lat, gpx = [], {}
    for k in range(C1_FRAMES):
        cam = G.camera(20.0, 10.0 * k, 3.0)
        view = G.make_view(C1_SIZE, C1_SIZE, cam, alpha_stop=ALPHA_STOP, precision=32)
        G.synchronize()
        t0 = time.perf_counter()
        frame.reset()
        G.render(cache, view, d_tf, len(tf), frame, d_st)
        G.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
        if k < cpu_frames:
            gpx[k] = frame.pixels().reshape(-1).copy()
    samples = int(d_st.to_array(np.uint64, (4,))[0])
    frame.close()
    cache.close()
    for b in d_levels:
        b.free()
    gpu = {"mip_ms": gpu_mip_ms, "brick_cut_ms": gpu_brick_ms, "load_ms": load_ms,
           "frames_per_s": C1_FRAMES / (sum(lat) / 1e3), "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
           "samples_per_s": samples / (sum(lat) / 1e3)}
    # CPU: preprocessor and FP64 caster
    levels, t_mip, t_brick, bricks = cpu_mip_and_bricks(O, vol, nl, C1_BRICK, 1, cores)
    ocfg = O.make_cfg(vc.initial(), vc.real(), C1_BRICK, 1, 1, cs)
    oc = O.Cache(ocfg, levels)
    oc.load_all()
    errs, csamp = [], 0
    t0 = time.perf_counter()
    for k in range(cpu_frames):
        view = O.make_view(ocfg, C1_SIZE, C1_SIZE, G.camera(20.0, 10.0 * k, 3.0), tf, alpha_stop=ALPHA_STOP)
        px = np.zeros(C1_SIZE * C1_SIZE, np.uint32)
        st = oc.render(view, np.zeros(C1_SIZE * C1_SIZE, np.float32), np.zeros(C1_SIZE * C1_SIZE * 4), px,
                       precision=64, nthreads=cores)
        csamp += int(st[0])
        errs.append(int(np.abs(G.unpack_rgba(px) - G.unpack_rgba(gpx[k])).max()))
    t_cast = time.perf_counter() - t0
    oc.close()
    mip_vox = sum(l.size for l in levels[1:])
    cpu = {"cores": cores, "mip_ms": t_mip * 1e3, "bricker_ms": t_brick * 1e3, "bricks": bricks,
           "frames_per_s": cpu_frames / t_cast, "samples_per_s": csamp / t_cast, "frames_timed": cpu_frames,
           "mip_voxels_per_s": mip_vox / t_mip, "bricker_bricks_per_s": bricks / t_brick}
    print("plain config: GPU %.1f frames/s (mip %.1f ms, cut %.1f ms, load %.1f ms); CPU %d cores %.2f frames/s "
          "(mip %.0f ms, bricker %.0f ms); max RGBA diff %s" % (gpu["frames_per_s"], gpu_mip_ms, gpu_brick_ms, load_ms,
                                                                cores, cpu["frames_per_s"], cpu["mip_ms"],
                                                                cpu["bricker_ms"], errs), file=sys.stderr, flush=True)
    return {"workload": "256^3 u8, 12 blobs + noise, 16^3 bricks, covering 1, 5 levels, resident; 512x512, 36-frame "
                        "orbit, alpha stop 0.99; GPU FP32 compositing, CPU FP64 caster on %d frames" % cpu_frames,
            "gpu": gpu, "cpu": cpu, "max_rgba_diff_gpu_vs_cpu": max(errs) if errs else None}

# batching workload: 64 uneven views of one 1024^3 u8 volume
BATCH_DIMS = (1024, 1024, 1024)
BATCH_BRICK = 16
BATCH_VIEWS = 64
BATCH_SIZES = [(256, 256), (640, 480), (1024, 1024), (1920, 1080)]


# serving workload: Poisson arrivals over 6 resident volumes
SERVE_VOLUMES = [((512, 512, 512), 1), ((1024, 1024, 1024), 1), ((1024, 1024, 1024), 2), ((2048, 2048, 2048), 1),
                 ((2048, 2048, 1024), 2), ((4096, 4096, 4096), 1)]
SERVE_POOL = 12 << 30          # per volume; 6 pools + view slots within the 96 GiB budget
SERVE_BUDGET = 96 << 30


def serving_workload(G, rates=(4, 8, 16, 32, 64), seconds=4.0, max_views=4, warm_requests=24, fixed_ab=False,
                     volumes=None, max_requests=16384, device_source=False, count=0, tune_upload=False):
    """Requests arrive as a Poisson process (seeded) over 6 volumes (512^3 u8,
    1024^3 u8, 1024^3 u16, 2048^3 u8, 2048x2048x1024 u16, 4096^3 u8, each
    synthesised brick by brick on the host on demand, 32^3 bricks, 12 GiB
    pool), each with its own square resolution (256..2048), camera and one of
    8 transfer functions.  For every rate the same arrival sequence is served
    twice: by the serving loop (continuous batching, up to `max_views` views
    per volume per launch, the six volumes' launches overlapping on their own
    streams) and by the plain renderer (one request at a time, one view per
    launch).  Each cache loads up to `max_requests` bricks per update (the
    upload batch; 16384 measured best against 4096 for both loops).  Latency is
    submission to pixels on the host.  With `device_source`, the bricks are
    synthesised on the GPU straight into each cache's staging buffer (same
    bytes, no host arithmetic, no PCIe upload).  With `count`, every rate
    serves that many requests (arrivals over count / rate seconds) instead of
    `seconds` of arrivals, which bounds the run at rates the GPU cannot
    sustain.  Requests resume stopped rays by exact continuation, so every
    result equals its one-pass render."""
    import threading

    tables = [G.tf_samples(G.tf_seeded(SEED + 200 + i, 6), 256) for i in range(8)]
    sources, vcs = [], []
    for i, (dims, nb) in enumerate(SERVE_VOLUMES):
        vc = G.volume_cfg(dims, 32, 1, nb)
        amp = (60.0, 255.0) if nb == 1 else (15000.0, 65535.0)
        params = G.synth_params(SEED + 10 + i, 32, sigma=(dims[0] / 40.0, dims[0] / 8.0), amp=amp,
                                noise=8 if nb == 1 else 2048, threads=0)
        sources.append(G.Source.synthetic(vc, params, device=0 if device_source else None))
        vcs.append(vc)
    d_tf0 = G.DeviceBuffer.from_array(tables[0])

    upload_table = None
    if tune_upload:
        # the upload batch, autotuned on cold frames of the 1024^3 u8 volume
        import autotune

        def make_cache():
            vc1 = vcs[1]
            cc = G.cache_cfg(G.pool_shape(SERVE_POOL, 32 ** 3), 32, max_requests=16384, sync=False)
            return G.Cache(vc1, cc, sources[1])

        fr = [(G.make_view(side, side, G.camera(15.0, a, 2.6), alpha_stop=ALPHA_STOP), d_tf0, 256)
              for side, a in ((1024, 30.0), (512, 160.0))]
        up = autotune.tune_upload(make_cache, fr, candidates=(4096, 16384))
        max_requests, upload_table = up["max_requests"], up["table"]
        print("serving: autotuned upload batch %d bricks" % max_requests, file=sys.stderr, flush=True)

    def make_requests(rate, seed, n=0):
        rng = np.random.default_rng(seed)
        t, out = 0.0, []
        while True:
            t += float(rng.exponential(1.0 / rate))
            if (n and len(out) >= n) or (not n and t >= seconds):
                return out
            side = int(rng.integers(4, 33)) * 64
            cam = G.camera(float(rng.uniform(-30.0, 30.0)), float(rng.uniform(0.0, 360.0)),
                           float(rng.uniform(2.2, 3.6)))
            vol = int(rng.integers(len(SERVE_VOLUMES)))
            if volumes is not None:
                vol = volumes[vol % len(volumes)]
            out.append((t, vol, side, cam, int(rng.integers(len(tables)))))

    def serve(reqs, plain, fixed=False):
        srv = G.Server(SERVE_BUDGET, max_views=1 if plain else max_views, slot_pixels=2048 * 2048, precision=32,
                       plain=plain, fixed_views=fixed)
        vids = [srv.add_volume(vc, src, SERVE_POOL, max_requests=max_requests) for vc, src in zip(vcs, sources)]
        # warm: an untimed burst from the same request distribution
        for (_, vol, side, cam, tfi) in warm:
            srv.submit(vids[vol], G.make_view(side, side, cam, alpha_stop=ALPHA_STOP, resume=1), tables[tfi])
        srv.drain()
        while True:
            rs = srv.poll(256)
            if not rs:
                break
            for r in rs:
                srv.take_pixels(r.id, r.width, r.height)
        st0 = srv.stats()
        s0 = st0.samples
        sizes, lat = {}, []
        t_start = time.perf_counter()

        def feeder():
            for (t, vol, side, cam, tfi) in reqs:
                d = t_start + t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                rid = srv.submit(vids[vol], G.make_view(side, side, cam, alpha_stop=ALPHA_STOP, resume=1), tables[tfi])
                sizes[rid] = (side, side)

        th = threading.Thread(target=feeder)
        th.start()
        done = 0
        while done < len(reqs):
            for r in srv.poll(256):
                srv.take_pixels(r.id, r.width, r.height)
                lat.append(r.latency_ms)
                done += 1
            time.sleep(0.001)
        th.join()
        wall = time.perf_counter() - t_start
        st1 = srv.stats()
        samples = st1.samples - s0
        srv.close()
        return {"completed": done, "wall_s": wall, "frames_per_s": done / wall, "samples_per_s": samples / wall,
                "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
                "bricks_loaded": st1.bricks_loaded - st0.bricks_loaded,
                "pressure_events": st1.pressure_events - st0.pressure_events,
                "scheduler": {"iterations": st1.iterations - st0.iterations,
                              "idle_iterations": st1.idle_iterations - st0.idle_iterations,
                              "launch_ms": st1.launch_ms - st0.launch_ms, "update_ms": st1.update_ms - st0.update_ms,
                              "fetch_ms": st1.fetch_ms - st0.fetch_ms},
                "ray_starts": st1.ray_starts - st0.ray_starts}

    warm = []
    while len(warm) < warm_requests:
        warm += make_requests(8.0, SEED + 299 + len(warm))
    warm = warm[:warm_requests]
    # the warm burst reaches every volume, so that no cache is created inside
    # the timed window of either loop
    seen = {w[1] for w in warm}
    for vol in range(len(SERVE_VOLUMES)):
        if vol not in seen and (volumes is None or vol in volumes):
            warm.append((0.0, vol, 256, G.camera(15.0, 30.0, 3.0), 0))
    out = []
    for i, rate in enumerate(rates):
        reqs = make_requests(rate, SEED + 300 + i, count)
        b = serve(reqs, False)
        bf = serve(reqs, False, fixed=True) if fixed_ab else None
        pl = serve(reqs, True)
        print("serving %d req/s: %d requests; batched p50 %.0f p99 %.0f ms %.1f fps %d bricks %d pressure | plain p50 "
              "%.0f p99 %.0f ms %.1f fps %d bricks" % (rate, len(reqs), b["p50_ms"], b["p99_ms"], b["frames_per_s"],
                                                       b["bricks_loaded"], b["pressure_events"], pl["p50_ms"],
                                                       pl["p99_ms"], pl["frames_per_s"], pl["bricks_loaded"]),
              file=sys.stderr, flush=True)
        if bf:
            print("serving %d req/s: batched, fixed %d views: p50 %.0f p99 %.0f ms %.1f fps %d bricks"
                  % (rate, max_views, bf["p50_ms"], bf["p99_ms"], bf["frames_per_s"], bf["bricks_loaded"]),
                  file=sys.stderr, flush=True)
        out.append({"rate": rate, "requests": len(reqs), "batched": b, "plain": pl})
        if bf:
            out[-1]["batched_fixed_views"] = bf
    arrivals = ("%d requests per rate" % count) if count else ("Poisson arrivals for %.0f s per rate" % seconds)
    res = {"workload": "%s over 6 volumes synthesised on demand %s(512^3 u8, 1024^3 u8, 1024^3 u16, 2048^3 u8, "
                       "2048x2048x1024 u16, 4096^3 u8; 32^3 bricks, 12 GiB pool each within one 96 GiB device budget "
                       "that admits a volume's pool while it fits), square views 256..2048, own camera and TF per "
                       "request; batched = serving loop (<= %d views per volume per launch), plain = one request at "
                       "a time; %d bricks per update" % (arrivals, "on the GPU " if device_source else "on the host ",
                                                         max_views, max_requests), "rates": out}
    # frames/s of each loop at equal p99: the batched loop's throughput at the
    # highest rate whose p99 does not exceed the plain loop's p99 at that rate
    ratios = [r["batched"]["frames_per_s"] / r["plain"]["frames_per_s"] for r in out if r["plain"]["frames_per_s"] > 0]
    res["batched_over_plain_frames_per_s"] = ratios
    if upload_table:
        res["upload_autotune"] = upload_table
    eq = [r for r in out if r["batched"]["p99_ms"] <= r["plain"]["p99_ms"]]
    if eq:
        best = max(eq, key=lambda r: r["batched"]["frames_per_s"] / r["plain"]["frames_per_s"])
        res["at_equal_or_lower_p99"] = {"rate": best["rate"],
                                        "speedup": best["batched"]["frames_per_s"] / best["plain"]["frames_per_s"]}
    return res


# on-demand processing: a virtual volume larger than HBM, coarse levels built
# on the GPU on miss
ONDEMAND_DIMS = (8192, 8192, 4096)
ONDEMAND_POOL = 128 << 30
# one page-table cache: 16384 blocks of 8x8x8 bricks (CacheSize[1], BlockSize[1])
ONDEMAND_TABLES = [((32, 32, 16), (8, 8, 8))]


def _lod_span(vc, width, dist):
    """Levels a centred view at `dist` samples (compute_LOD, RayCasters.cu:
    551-567): the level at the volume's front face and at its back, from the
    pixel footprint dist_along_ray / width (the view's half-angle is
    atan(1/2) over width/2 pixels)."""
    def level_at(t):
        pd = t / width
        lev = 0
        for l in range(1, vc.nb_levels):
            sp = [2.0 / vc.initial_sizes[l][a] for a in range(3)]
            if max(sp[0], sp[1]) < pd:
                lev = l
            else:
                break
        return lev
    return level_at(max(dist - 1.0, 1e-6)), level_at(dist + 1.0)


def directory_bytes(g):
    """Device bytes of the page directory a cache keeps besides its request
    buffer (which the reference sizes by the bricks either way): the MRPD
    (uint4 per cell) and the ray caster's cell words, plus per table cache
    its entry words and per-block usage, LRU order (two copies) and owner."""
    b = int(g.top_cells) * (16 + 4)
    for i in range(int(g.nb_tables)):
        t = g.table[i]
        b3 = int(t.block[0]) * int(t.block[1]) * int(t.block[2])
        b += int(t.nblocks) * (b3 * 4 + 4 + 8 + 16)
    return b


def ondemand_workload(G, sessions=32, steps=9, width=1280, height=720, far=200.0, near=1.05, tables=()):
    """The on-demand configuration: a virtual 8192x8192x4096 u8 volume
    (256 GiB at level 0, 9 levels, 32^3 bricks, covering 1, 128 GiB pool).
    Level 0 is synthesised on the GPU from seeded hash noise + blobs when a
    brick is requested; the coarser levels are its mip pyramid, built once on
    the GPU (cold build, timed apart: level 0 synthesised and reduced in
    z-slabs, levels 1-8 kept resident) and cut into bricks on request.
    `sessions` interactive sessions of 1280x720 views, each its own orbit
    angle and transfer function, zoom from `far` to `near` (distance to the
    volume centre) in `steps` geometric steps, so the path spans the coarsest
    to the finest level; each step renders the sessions' views as one batch."""
    vc = G.volume_cfg(ONDEMAND_DIMS, 32, 1, 1)
    params = G.synth_params(SEED + 5, 48, sigma=(100.0, 900.0), amp=(60.0, 255.0), noise=8, threads=0)
    src, build_ms = G.Source.gpu_pyramid(vc, params, slab_planes=64)
    print("ondemand: cold pyramid build %.0f ms (levels 1-%d resident)" % (build_ms, vc.nb_levels - 1),
          file=sys.stderr, flush=True)
    cs = G.pool_shape(ONDEMAND_POOL, 32 ** 3)
    cache = G.Cache(vc, G.cache_cfg(cs, 32, max_requests=16384, sync=False, tables=tables), src)
    flat_dir = directory_bytes(G.geometry_compute(vc, G.cache_cfg(cs, 32)))
    dir_bytes = directory_bytes(cache.g)
    tfs = [G.tf_samples(G.tf_seeded(SEED + 400 + i, 6), 256) for i in range(sessions)]
    d_tfs = [G.DeviceBuffer.from_array(t) for t in tfs]
    b = G.Batch(cache, sessions, sessions * width * height, 32)
    out = []
    for k in range(steps):
        dist = far * (near / far) ** (k / max(1, steps - 1))
        views = [G.make_view(width, height, G.camera(15.0 + 10.0 * (i % 3), 45.0 * i, dist), alpha_stop=ALPHA_STOP,
                             precision=32) for i in range(sessions)]
        c0 = cache.counters()
        t0 = time.perf_counter()
        r = b.render(views, d_tfs, [256] * sessions)
        dt = time.perf_counter() - t0
        c1 = cache.counters()
        lru = cache.lru()   # (slot origin xyz, level, volume position xyz as float bits; -1: free)
        levels = lru[:, 3][lru[:, 4] != np.float32(-1.0).view(np.uint32)]
        span = _lod_span(vc, width, dist)
        step = {"distance": dist, "lod_front_back": list(span), "ms": dt * 1e3, "frames_per_s": sessions / dt,
                "launches": r.launches, "bricks_loaded": c1["bricks_loaded"] - c0["bricks_loaded"],
                "fetch_ms": c1["fetch_ms"] - c0["fetch_ms"], "samples": r.stats.mapped_samples,
                "complete": bool(r.complete)}
        step["resident_bricks_per_level"] = np.bincount(levels, minlength=vc.nb_levels).tolist()
        if tables:
            step["unreferenced_bricks_last_load"] = cache.unreferenced()
        print("ondemand step %d distance %.2f (levels %d..%d): %.0f ms, %d launches, %d bricks, fetch %.0f ms, "
              "%.3g samples" % (k, dist, span[0], span[1], dt * 1e3, r.launches, step["bricks_loaded"],
                                step["fetch_ms"], step["samples"]), file=sys.stderr, flush=True)
        out.append(step)
    b.close()
    cache.close()
    src.close()
    lat = [s["ms"] for s in out]
    tdesc = ""
    if tables:
        tdesc = "; page-table cache%s %s" % ("s" if len(tables) > 1 else "", ", ".join(
            "%s blocks of %s bricks" % ("x".join(map(str, c)), "x".join(map(str, b))) for c, b in tables))
    return {"workload": "virtual 8192x8192x4096 u8 (9 levels, 32^3 bricks, covering 1, 128 GiB pool); level 0 "
                        "synthesised on the GPU on request, levels 1-8 its mip pyramid built cold on the GPU; %d "
                        "sessions of %dx%d zooming %.0f -> %.2f in %d steps (levels 8 -> 0), one batch per step%s"
                        % (sessions, width, height, far, near, steps, tdesc),
            "directory_bytes": dir_bytes, "directory_bytes_without_tables": flat_dir,
            "cold_pyramid_build_ms": build_ms,
            "frames_per_s": sessions * len(out) / (sum(lat) / 1e3), "step_p50_ms": percentile(lat, 50),
            "step_p99_ms": percentile(lat, 99), "steps": out}


def batch_workload(G, reps=3):
    """Batched vs plain rendering of 64 concurrent views (sizes uniform among
    256^2, 640x480, 1024^2, 1920x1080, seeded cameras, 8 seeded 256-entry
    transfer functions) over one 1024^3 u8 blob volume synthesised on demand
    (16^3 bricks, covering 1, 4 GiB pool).  The autotuner first picks the
    upload batch (bricks per update) on cold renders of 8 of the views, then
    the batch is rendered cold (loading its bricks), then the autotuner picks
    the ray tile (8x8, 16x8, 32x4, 64x1) x samples-per-check (1-4) on the
    warm batch, then: batched = one gc_batch_render per batch (hipGraph
    replay), plain = gc_render_frame per view, one view per launch, in request
    order.  All 64 requests arrive together, so a view's latency is the time
    until its own frame is ready."""
    import autotune

    rng = np.random.default_rng(SEED + 3)
    vc = G.volume_cfg(BATCH_DIMS, BATCH_BRICK, 1, 1)
    params = G.synth_params(SEED + 3, 32, sigma=(10.0, 60.0), amp=(60.0, 255.0), noise=8, threads=0)
    src = G.Source.synthetic(vc, params)
    cs = G.pool_shape(4 << 30, BATCH_BRICK ** 3)

    def make_cache():
        return G.Cache(vc, G.cache_cfg(cs, BATCH_BRICK, max_requests=16384, sync=False), src)

    tables = [G.tf_samples(G.tf_seeded(SEED + 100 + i, 6), 256) for i in range(8)]
    d_tfs = [G.DeviceBuffer.from_array(t) for t in tables]
    views, vt = [], []
    for i in range(BATCH_VIEWS):
        w, h = BATCH_SIZES[int(rng.integers(len(BATCH_SIZES)))]
        cam = G.camera(float(rng.uniform(-40.0, 40.0)), float(rng.uniform(0.0, 360.0)), float(rng.uniform(2.2, 3.6)))
        views.append(G.make_view(w, h, cam, alpha_stop=ALPHA_STOP, precision=32))
        vt.append(d_tfs[i % len(d_tfs)])
    sizes = [256] * BATCH_VIEWS
    npix = sum(v.width * v.height for v in views)
    upload = autotune.tune_upload(make_cache, [(views[i], vt[i], 256) for i in range(0, BATCH_VIEWS, 8)],
                                  candidates=(2048, 4096, 8192, 16384))
    print("batch: autotuned upload batch %d bricks" % upload["max_requests"], file=sys.stderr, flush=True)
    cache = make_cache()
    G.set_max_requests(cache, upload["max_requests"])
    b = G.Batch(cache, BATCH_VIEWS, npix, 32)
    t0 = time.perf_counter()
    cold = b.render(views, vt, sizes)
    cold_ms = (time.perf_counter() - t0) * 1e3
    print("batch: cold render %.0f ms, %d launches, %d bricks" % (cold_ms, cold.launches, cold.bricks_loaded),
          file=sys.stderr, flush=True)
    tune = autotune.tune_launch(cache, views, vt, sizes, tiles=(0, 4, 5, 3), marches=(1, 2, 3, 4), reps=2)
    print("batch: autotuned tile %d march %d" % (tune["tile"], tune["march"]), file=sys.stderr, flush=True)
    for v in views:
        v.tile, v.march = tune["tile"], tune["march"]
    b.set_graph(True)
    b.render(views, vt, sizes)                      # capture
    bt, bs = [], 0
    for _ in range(reps):
        t0 = time.perf_counter()
        r = b.render(views, vt, sizes)
        bt.append((time.perf_counter() - t0) * 1e3)
        bs = r.stats.mapped_samples
    b.close()
    frames = [G.Frame(v.width, v.height, 32) for v in views]
    pt, plat, ps = [], [], 0
    for _ in range(reps):
        t0 = time.perf_counter()
        lat = []
        ps = 0
        for v, d, n, f in zip(views, vt, sizes, frames):
            f.reset()
            r = G.render_frame(cache, v, d, n, f)
            ps += r.stats.mapped_samples
            lat.append((time.perf_counter() - t0) * 1e3)
        pt.append(lat[-1])
        plat.extend(lat)
    for f in frames:
        f.close()
    cache.close()
    bmed, pmed = float(np.median(bt)), float(np.median(pt))
    return {
        "workload": "64 uneven views (256^2/640x480/1024^2/1920x1080), 8 TFs, 1024^3 u8 blobs, 16^3 bricks, "
                    "4 GiB pool, warm; batched = one hipGraph-replayed launch per batch, plain = one view per launch",
        "views": BATCH_VIEWS, "pixels": npix,
        "batched_frames_per_s": BATCH_VIEWS / (bmed / 1e3), "plain_frames_per_s": BATCH_VIEWS / (pmed / 1e3),
        "speedup": pmed / bmed,
        "batched_samples_per_s": bs / (bmed / 1e3), "plain_samples_per_s": ps / (pmed / 1e3),
        "batched_p50_ms": bmed, "batched_p99_ms": percentile(bt, 99),
        "plain_p50_ms": percentile(plat, 50), "plain_p99_ms": percentile(plat, 99),
        "tile": TILE_NAMES[tune["tile"]], "march": tune["march"], "upload_batch": upload["max_requests"],
        "upload_autotune": upload["table"],
        "autotune": [{"tile": t["tile"], "march": t["march"], "kernel_ms": round(t["kernel_ms"], 2)}
                     for t in tune["table"]],
        "cold_ms": cold_ms,
    }


def _pmc_pass(counters, outdir, args):
    """One child run of this benchmark under rocprofv3 collecting
    `counters` (one rocprofv3 pass); returns ({counter: per-k_dvr-dispatch
    values after the first frame}, {counter: copy kernel value})."""
    tag = counters[0]
    cmd = ["rocprofv3", "--pmc"] + list(counters) + ["-d", outdir, "-o", tag, "--", sys.executable,
           os.path.abspath(__file__), "--steps", str(args.steps), "--warmup", str(args.warmup), "--no-cpu-baseline",
           "--no-traffic", "--no-extras", "--calib-copy-mib", str(CALIB_MIB), "--max-requests", str(args.max_requests),
           "--tile", str(args.tile), "--precision", str(args.precision)] + \
          (["--handoff", args.handoff] if args.handoff else [])
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError("rocprofv3 %s pass failed (%d): %s" % (tag, r.returncode, r.stderr[-400:]))
    first_launches = None   # k_dvr launches of the warm-up frames (the timed frames' follow)
    for line in r.stderr.splitlines():
        if line.startswith("frame ") and " warmup " in line:
            first_launches = (first_launches or 0) + int(line.split(" lane ")[1].split()[0])
    dbs = [os.path.join(dp, f) for dp, _, fs in os.walk(outdir) for f in fs if f.endswith(".db") and tag in f]
    if not dbs or first_launches is None:
        raise RuntimeError("rocprofv3 %s pass left no results" % tag)
    con = sqlite3.connect(dbs[0])
    vals, copies = {}, {}
    for counter in counters:
        rows = con.execute("select dispatch_id, kernel_name, sum(value) from counters_collection where counter_name = ? "
                           "group by dispatch_id order by dispatch_id", (counter,)).fetchall()
        # a lane launch's hand-over continuation (a k_dvr_ray dispatch right
        # after its k_dvr dispatch) counts with it, as in its HIP-event time
        dvr = []
        for i, (_, k, v) in enumerate(rows):
            if "k_dvr<" in k:
                dvr.append(v)
            elif "k_dvr_ray<" in k and i > 0 and "k_dvr<" in rows[i - 1][1] and dvr:
                dvr[-1] += v
        copy = [v for _, k, v in rows if "copy" in k.lower()]
        if len(dvr) <= first_launches or not copy:
            con.close()
            raise RuntimeError("rocprofv3 %s pass: no ray-caster or copy dispatches" % counter)
        vals[counter] = dvr[first_launches:]
        copies[counter] = max(copy)
    con.close()
    return vals, copies


def measure_traffic(args):
    """HBM bytes per ray-caster launch from FETCH_SIZE + WRITE_SIZE (KiB per
    dispatch), each scaled by (bytes the calibration copy moves) / (what the
    counter reads for it), and the ray caster's L2 hit rate
    TCC_HIT / (TCC_HIT + TCC_MISS) over the same launches.  Three rocprofv3
    passes, one counter group each."""
    if not shutil.which("rocprofv3"):
        return None, "rocprofv3 not found", None
    tmp = tempfile.mkdtemp(prefix="gcvol_pmc_", dir="/tmp")
    try:
        out = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            t0 = time.perf_counter()
            vals, copies = _pmc_pass([counter], os.path.join(tmp, counter), args)
            v, copy_kib = vals[counter], copies[counter]
            scale = (CALIB_MIB * 1024.0) / copy_kib
            print("traffic pass %s: %d launches, copy %.0f KiB, %.1f s" % (counter, len(v), copy_kib,
                                                                           time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
            out[counter] = (sum(v) / len(v) * scale * 1024.0, scale, len(v))
        total = out["FETCH_SIZE"][0] + out["WRITE_SIZE"][0]
        note = ("FETCH_SIZE x%.3f + WRITE_SIZE x%.3f (copy calibration, %d MiB), mean over the %d k_dvr launches of "
                "the timed frames" % (out["FETCH_SIZE"][1], out["WRITE_SIZE"][1], CALIB_MIB, out["FETCH_SIZE"][2]))
        t0 = time.perf_counter()
        vals, _ = _pmc_pass(["TCC_HIT_sum", "TCC_MISS_sum"], os.path.join(tmp, "TCC"), args)
        hit, miss = sum(vals["TCC_HIT_sum"]), sum(vals["TCC_MISS_sum"])
        l2 = {"hit_rate": hit / (hit + miss) if hit + miss > 0 else None,
              "hits_per_launch": hit / len(vals["TCC_HIT_sum"]), "misses_per_launch": miss / len(vals["TCC_MISS_sum"]),
              "note": "TCC_HIT_sum / (TCC_HIT_sum + TCC_MISS_sum) over the k_dvr launches of the timed frames"}
        print("traffic pass TCC: %d launches, L2 hit rate %.3f, %.1f s" % (len(vals["TCC_HIT_sum"]), l2["hit_rate"] or 0,
                                                                         time.perf_counter() - t0),
              file=sys.stderr, flush=True)
        return total, note, l2
    except (RuntimeError, subprocess.TimeoutExpired, sqlite3.Error) as e:
        return None, str(e)[:300], None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--max-requests", type=int, default=0,
                    help="bricks per update (0: picked by the autotuner on cold frames of the path)")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--precision", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 counter passes")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain workload")
    ap.add_argument("--no-serving", action="store_true", help="skip the bounded serving leg")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline workload (no plain-config, batch or serving legs)")
    ap.add_argument("--serving", action="store_true",
                    help="also run the long serving workload (host-synthesised bricks, several minutes)")
    ap.add_argument("--serving-only", action="store_true", help="run only the serving workload")
    ap.add_argument("--no-ondemand", action="store_true", help="skip the on-demand 8192x8192x4096 workload")
    ap.add_argument("--ondemand-only", action="store_true",
                    help="run only the on-demand workload (virtual 8192x8192x4096 u8, GPU pyramid)")
    ap.add_argument("--serving-rates", default="4,8,16,32,64", help="request rates (req/s) of the serving workload")
    ap.add_argument("--serving-volumes", default="", help="serve only these volumes (indices into the six)")
    ap.add_argument("--serving-max-views", type=int, default=4, help="views per volume per launch when batched")
    ap.add_argument("--serving-max-requests", type=int, default=16384, help="bricks per update of each served volume")
    ap.add_argument("--serving-source", choices=("host", "gpu"), default="host",
                    help="where the served volumes' bricks are synthesised")
    ap.add_argument("--serving-ab", action="store_true",
                    help="also serve each arrival sequence with a fixed number of views per volume")
    ap.add_argument("--handoff", default="",
                    help="SAMPLES[:LANES[:PCT]]: lane-per-ray launches hand rays that have consumed SAMPLES samples "
                         "over to a group march of LANES lanes per ray within the launch once PCT %% of its waves have "
                         "run out of rays (0: off; default: the library's setting)")
    ap.add_argument("--launch-log", action="store_true",
                    help="log every ray-caster launch of the timed frames (stderr) and add the last frame's to the "
                         "output line")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--pool-gib", type=float, default=0.0, help="brick pool size (default: the workload's 4 GiB)")
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
    if args.ondemand_only:
        print(json.dumps({"ondemand": ondemand_workload(G), "ondemand_tables": ondemand_workload(G, tables=ONDEMAND_TABLES)}))
        return
    if args.serving_only:
        rates = tuple(int(r) for r in args.serving_rates.split(","))
        vols = [int(v) for v in args.serving_volumes.split(",")] if args.serving_volumes else None
        print(json.dumps({"serving": serving_workload(G, rates=rates, fixed_ab=args.serving_ab, volumes=vols,
                                                      max_views=args.serving_max_views,
                                                      max_requests=args.serving_max_requests,
                                                      device_source=args.serving_source == "gpu")}))
        return

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
    cs = G.pool_shape(POOL_BYTES if args.pool_gib <= 0 else int(args.pool_gib * (1 << 30)), brick_bytes)
    tf = G.tf_samples(G.tf_seeded(SEED, 6), 256)
    d_tf = G.DeviceBuffer.from_array(tf)
    upload_tune = None
    hs = args.handoff.split(":") if args.handoff else []
    handoff = tuple(int(hs[i]) if len(hs) > i else 0 for i in range(3)) if hs else None
    if args.max_requests <= 0:
        # the upload batch, picked by the autotuner on three cold frames of
        # the path (fresh caches), before anything is timed
        import autotune

        def make_cache():
            c = G.Cache(vc, G.cache_cfg(cs, BRICK, max_requests=16384, sync=False), src)
            if handoff is not None:
                G.set_ray_handoff(c, *handoff)
            return c

        fr = [(G.make_view(WIDTH, HEIGHT, fly_through(pf), alpha_stop=ALPHA_STOP, precision=args.precision,
                           tile=args.tile), d_tf, len(tf)) for pf in (0, 54, 108)]
        upload_tune = autotune.tune_upload(make_cache, fr, candidates=(2048, 4096, 8192, 16384))
        args.max_requests = upload_tune["max_requests"]
        print("autotuned upload batch: %d bricks per update %s" % (args.max_requests, upload_tune["table"]),
              file=sys.stderr, flush=True)
    # counters of the timed frames' ray-caster launches, from child runs of
    # this benchmark with the same frames and upload batch
    traffic, traffic_note, l2 = (None, "not measured", None) if args.no_traffic else measure_traffic(args)
    cc = G.cache_cfg(cs, BRICK, max_requests=args.max_requests, sync=False)
    cache = G.Cache(vc, cc, src)
    if handoff is not None:
        G.set_ray_handoff(cache, *handoff)
    handoff = G.get_ray_handoff(cache)
    frame = G.Frame(WIDTH, HEIGHT, args.precision)
    log = G.LaunchLog(cache) if args.launch_log else None
    last_log = None

    nframes = args.warmup + args.steps
    lat, reports = [], []
    cams = []
    for k in range(nframes):
        cam = fly_through(path_frame(k, args.warmup, args.steps))
        cams.append(cam)
        view = G.make_view(WIDTH, HEIGHT, cam, alpha_stop=ALPHA_STOP, precision=args.precision, tile=args.tile)
        G.synchronize()
        fetch0 = cache.counters()["fetch_ms"]
        t0 = time.perf_counter()
        frame.reset()
        rep = G.render_frame(cache, view, d_tf, len(tf), frame)
        dt = time.perf_counter() - t0
        d = rep.as_dict()
        d["wall_ms"] = dt * 1e3
        d["fetch_ms"] = cache.counters()["fetch_ms"] - fetch0
        print("frame %3d path %3d %s wall %.1f ms launches %d lane %d bricks %d samples %.3g rays %d kernel %.1f ms "
              "fetch %.1f ms" % (k, path_frame(k, args.warmup, args.steps), "timed" if k >= args.warmup else "warmup",
                                 dt * 1e3, d["launches"],
                                 d["lane_launches"], d["bricks_loaded"], d["mapped_samples"], d["pixels"],
                                 d["kernel_ms"], d["fetch_ms"]), file=sys.stderr, flush=True)
        if k >= args.warmup:
            lat.append(dt * 1e3)
            reports.append(d)
            if log is not None:
                last_log = log.last_frame()
                for r in last_log:
                    print("  launch %3d %-5s list %8d group %2d rays %8d samples %11d visits %9d stopped %8d "
                          "bricks %6d %.3f ms handed %d" % (r["launch"], r["kernel"], r["list_rays"], r["group"],
                                                            r["rays"], r["mapped_samples"], r["brick_visits"],
                                                            r["unmapped_rays"], r["bricks_loaded"], r["kernel_ms"],
                                                            r["handed_over"]), file=sys.stderr, flush=True)
        if k == args.warmup and not args.no_cpu_baseline:
            # untimed: the CPU baseline over this frame's cache state
            cpu = cpu_baseline(G, cache, vc, cam, tf, seconds=args.cpu_seconds,
                               device_pixels=frame.pixels().reshape(-1))
    total_s = sum(lat) / 1e3
    S = sum(r["mapped_samples"] for r in reports)
    kms = sum(r["kernel_ms"] for r in reports)
    launches = sum(r["launches"] for r in reports)
    tot = {k: sum(r[k] for r in reports) for k in ("mapped_samples", "brick_visits", "pixels")}
    B = algorithmic_bytes(tot, BYTES)
    # the dominant kernel: k_dvr (tile / lane-per-ray launches), timed alone
    lane = {k: sum(r["lane_" + k] for r in reports) for k in ("mapped_samples", "brick_visits", "pixels")}
    lane_launches = sum(r["lane_launches"] for r in reports)
    lane_ms = sum(r["lane_kernel_ms"] for r in reports)
    B_lane = algorithmic_bytes(lane, BYTES)
    achieved = B_lane / (lane_ms / 1e3) if lane_ms > 0 else 0.0
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
            "tile": TILE_NAMES[args.tile],
            "upload_batch": args.max_requests,
            "upload_autotune": upload_tune["table"] if upload_tune else None,
            "handoff": {"samples": handoff[0], "lanes": handoff[1], "exited_pct": handoff[2]},
        },
        "frames_per_s": len(lat) / total_s if total_s > 0 else 0.0,
        "p50_ms": percentile(lat, 50),
        "p99_ms": percentile(lat, 99),
        "kernel_samples_per_s": S / (kms / 1e3) if kms > 0 else 0.0,
        "launches": launches,
        "bricks_loaded": sum(r["bricks_loaded"] for r in reports),
        "frame_split_ms": {"kernel": kms / len(reports) if reports else 0.0,
                           "brick_fetch": sum(r["fetch_ms"] for r in reports) / len(reports) if reports else 0.0,
                           "wall": sum(lat) / len(lat) if lat else 0.0},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "traffic_note": traffic_note,
            "l2": l2,
            "kernel": "k_dvr<2,%s,%s>" % ("double" if args.precision == 64 else "float",
                                          ["8,8", "16,4", "32,2", "64,1", "16,4", "32,2"][args.tile]) +
                      (" + hand-over k_dvr_ray<2,%s,%d,32>" % ("double" if args.precision == 64 else "float", handoff[1])
                       if handoff[0] else ""),
            "launches": lane_launches,
            "avg_launch_ms": lane_ms / lane_launches if lane_launches else 0.0,
            "algorithmic_bytes_per_launch": B_lane / lane_launches if lane_launches else 0.0,
            "all_ray_caster_launches": {
                "kernels": "k_dvr + k_dvr_ray", "launches": launches,
                "avg_launch_ms": kms / launches if launches else 0.0,
                "algorithmic_bytes_per_launch": B / launches if launches else 0.0,
                "achieved": (B / (kms / 1e3) if kms > 0 else 0.0) / 1e9,
            },
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu
    if log is not None:
        out["launch_log_last_frame"] = last_log
        log.close()
    cache.close()
    frame.close()
    if not args.no_extras:
        pc = plain_config_workload(G)
        out["plain_config"] = pc
        if "cpu_baseline" in out:
            c = pc["cpu"]
            out["cpu_baseline"]["preprocess"] = {
                "volume": "256^3 u8, 16^3 bricks, covering 1, 5 levels", "cores": c["cores"],
                "mip_voxels_per_s": c["mip_voxels_per_s"], "bricker_bricks_per_s": c["bricker_bricks_per_s"],
                "mip_ms": c["mip_ms"], "bricker_ms": c["bricker_ms"], "bricks": c["bricks"]}
        if not args.no_batch:
            out["batched_vs_plain"] = batch_workload(G)
        if not args.no_serving:
            out["serving"] = serving_workload(G, rates=(50, 200, 800), count=48, max_views=8, warm_requests=8,
                                              device_source=True, tune_upload=True)
        if not args.no_ondemand:
            out["ondemand"] = ondemand_workload(G)
            # the same through one page-table cache (MRPD of 8^3-brick blocks)
            out["ondemand_tables"] = ondemand_workload(G, tables=ONDEMAND_TABLES)
    if args.serving:
        out["serving_long"] = serving_workload(G)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
