"""Seeded test scenes shared by the parity tests (inputs only; no checking)."""
import numpy as np

import gcvol as G
import oracle_lib as O


def pyramid(vol, nb_levels):
    """Level 0 plus oracle mips (the CPU preprocessor) as [z,y,x] arrays."""
    levels = [np.ascontiguousarray(vol)]
    for _ in range(1, nb_levels):
        levels.append(O.mip_reduce(levels[-1]))
    return levels


def blob_volume(dims, bytes_, seed=99, nb_blobs=6, sigma=(4.0, 12.0), amp=None, noise=8):
    if amp is None:
        amp = (60.0, 255.0) if bytes_ == 1 else (15000.0, 65535.0)
    p = G.synth_params(seed, nb_blobs, sigma=sigma, amp=amp, noise=noise)
    return G.synth_volume(p, dims, bytes_)


class Scene:
    """A volume, its pyramid, matching oracle and product configurations."""

    def __init__(self, dims, bs, cov, bytes_, cache_size=None, max_requests=50, seed=99, big_brick=(1, 1, 1),
                 empty_threshold=0.0, tables=(), **blob_kw):
        self.dims, self.bs, self.cov, self.bytes = dims, bs, cov, bytes_
        self.vc = G.volume_cfg(dims, bs, cov, bytes_, big_brick)
        self.initial = self.vc.initial()
        self.real = self.vc.real()
        self.vol = blob_volume(dims, bytes_, seed=seed, **blob_kw)
        self.levels = pyramid(self.vol, self.vc.nb_levels)
        nb, _, _ = O.init_meta(self.initial, bs, cov, big_brick)
        self.total_bricks = int(sum(int(np.prod(x)) for x in nb))
        if cache_size is None:
            cache_size = O.pool_shape(self.total_bricks * bs ** 3 * bytes_, bs ** 3 * bytes_)
        self.cache_size = tuple(int(c) for c in cache_size)
        self.max_requests = max_requests
        self.big_brick = big_brick
        self.empty_threshold = empty_threshold
        self.tables = tuple(tables)
        self.ocfg = O.make_cfg(self.initial, self.real, bs, cov, bytes_, self.cache_size, bb=big_brick,
                               max_requests=max_requests, empty_threshold=empty_threshold, tables=self.tables)
        self.cc = G.cache_cfg(self.cache_size, bs, max_requests=max_requests, sync=True,
                              empty_threshold=empty_threshold, tables=self.tables)

    def oracle_cache(self):
        return O.Cache(self.ocfg, self.levels)

    def gpu_cache(self):
        src = G.Source.pyramid(self.vc, self.levels)
        c = G.Cache(self.vc, self.cc, src)
        c._src = src
        return c

    def resident_gpu_cache(self):
        """A cache whose pool holds every brick, fully loaded."""
        shape = O.pool_shape(self.total_bricks * self.bs ** 3 * self.bytes, self.bs ** 3 * self.bytes)
        src = G.Source.pyramid(self.vc, self.levels)
        c = G.Cache(self.vc, G.cache_cfg(shape, self.bs, max_requests=self.max_requests, sync=True,
                                         tables=self.tables), src)
        c._src = src
        c.load_all(self.device_levels())
        return c

    def device_levels(self):
        return [G.DeviceBuffer.from_array(l) for l in self.levels]


def unit_float(sc):
    """A float scene's volume scaled to [0, 1] (the transfer function's
    domain for float voxels), pyramid rebuilt; returns the scene."""
    assert sc.bytes == 4
    sc.vol = (sc.vol / np.float32(sc.vol.max())).astype(np.float32)
    sc.levels = pyramid(sc.vol, sc.vc.nb_levels)
    return sc


def tf_table(seed=5, n=256):
    return G.tf_samples(G.tf_seeded(seed, 6), n)


def orbit(k, rx=20.0, step=10.0, distance=3.0):
    return G.camera(rx, step * k, distance)
