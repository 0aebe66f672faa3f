"""ctypes binding of libgcvol.so (the C ABI declared in include/gcvol.h).

This module is host plumbing for tests and benchmarks: every computation
runs in the HIP library.  Importing it fails loudly if the library has not
been built (there is no fallback path).
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libgcvol.so")
MAXL = 16

u32 = C.c_uint32
u64 = C.c_uint64
f32 = C.c_float
vp = C.c_void_p

GC_OK = 0
DT = {1: np.uint8, 2: np.uint16, 4: np.float32}


class GcError(RuntimeError):
    pass


class VolumeCfg(C.Structure):
    _fields_ = [
        ("brick_size", u32 * 3), ("big_brick_size", u32 * 3), ("covering", u32 * 3),
        ("encoded_bytes", u32), ("channels", u32), ("nb_levels", u32),
        ("initial_sizes", (u32 * 3) * MAXL), ("real_sizes", (u32 * 3) * MAXL),
    ]

    def initial(self):
        return np.array([[self.initial_sizes[l][a] for a in range(3)] for l in range(self.nb_levels)], np.uint32)

    def real(self):
        return np.array([[self.real_sizes[l][a] for a in range(3)] for l in range(self.nb_levels)], np.uint32)


class CacheCfg(C.Structure):
    _fields_ = [
        ("cache_size", u32 * 3), ("block_size", u32 * 3), ("data_cache_flags", u32),
        ("max_requests", u32), ("sync_requests", u32), ("empty_threshold", f32),
    ]


class SynthParams(C.Structure):
    _fields_ = [
        ("seed", u64), ("nb_blobs", u32), ("sigma_min", f32), ("sigma_max", f32),
        ("amp_min", f32), ("amp_max", f32), ("noise", u32), ("threads", u32),
    ]


class Geometry(C.Structure):
    _fields_ = [
        ("nb_levels", u32), ("brick_size", u32), ("covering", u32), ("inner", u32), ("encoded_bytes", u32),
        ("level_dim", (u32 * 3) * MAXL), ("level_coord", u32 * MAXL), ("virt_vox", (u32 * 3) * MAXL),
        ("ior", (f32 * 3) * MAXL), ("rdv", (f32 * 3) * MAXL), ("one_over", f32 * 3),
        ("mrpd_dim", u32 * 3), ("cache_size", u32 * 3), ("nslots", u32), ("pad", u32), ("mrpd_cells", u64),
    ]


class View(C.Structure):
    _fields_ = [
        ("width", u32), ("height", u32), ("inv_mv", f32 * 12), ("bbox_min", f32 * 3), ("bbox_max", f32 * 3),
        ("alpha_stop", f32), ("precision", u32), ("tile", u32), ("reserved", u32),
    ]


class RenderStats(C.Structure):
    _fields_ = [("mapped_samples", u64), ("brick_visits", u64), ("pixels", u64), ("unmapped_rays", u64)]


class CacheView(C.Structure):
    _fields_ = [
        ("mrpd", vp), ("request", vp), ("usage", vp), ("pool", vp), ("geometry", vp),
        ("timestamp", u32), ("brick_size", u32), ("covering", u32), ("encoded_bytes", u32),
    ]


EXPORTS = [
    "gc_last_error", "gc_version", "gc_volume_cfg_init", "gc_pool_shape", "gc_cfg_get",
    "gc_device_count", "gc_set_device", "gc_malloc", "gc_free", "gc_memcpy_h2d", "gc_memcpy_d2h",
    "gc_memset", "gc_synchronize", "gc_stream_create", "gc_stream_destroy", "gc_stream_synchronize",
    "gc_mip_reduce", "gc_brick_cut", "gc_source_create_pyramid", "gc_source_create_synthetic",
    "gc_source_create_callback", "gc_source_destroy", "gc_source_get", "gc_synth_volume",
    "gc_cache_create", "gc_cache_destroy", "gc_cache_update", "gc_cache_completude", "gc_cache_load_all",
    "gc_cache_timestamp", "gc_cache_stream", "gc_cache_view_get", "gc_cache_geometry", "gc_cache_read_mrpd",
    "gc_cache_read_request", "gc_cache_write_request", "gc_cache_read_usage", "gc_cache_write_usage",
    "gc_cache_read_lru", "gc_cache_read_pool", "gc_cache_read_asked", "gc_frame_create", "gc_frame_destroy",
    "gc_frame_reset", "gc_frame_read_pixels", "gc_frame_read_state", "gc_frame_buffers", "gc_render_dvr",
    "gc_tf_samples", "gc_tf_seeded", "gc_camera_matrix",
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GcError("libgcvol.so is not built (run: python3 __graft_entry__.py build)")
        L = C.CDLL(LIB_PATH)
        for name in EXPORTS:
            getattr(L, name).restype = C.c_int
        L.gc_last_error.restype = C.c_char_p
        L.gc_version.restype = C.c_char_p
        L.gc_pool_shape.argtypes = [u64, u64, vp]
        L.gc_tf_seeded.argtypes = [u64, u32, vp]
        L.gc_camera_matrix.argtypes = [f32, f32, vp, vp]
        L.gc_render_dvr.argtypes = [vp, C.POINTER(View), vp, u32, vp, vp, vp]
        L.gc_frame_create.argtypes = [u32, u32, u32, C.POINTER(vp)]
        L.gc_mip_reduce.argtypes = [vp, vp, u32, vp, vp]
        L.gc_brick_cut.argtypes = [vp, vp, u32, u32, u32, vp, u32, vp, vp]
        L.gc_malloc.argtypes = [C.POINTER(vp), C.c_size_t]
        L.gc_free.argtypes = [vp]
        L.gc_memcpy_h2d.argtypes = [vp, vp, C.c_size_t]
        L.gc_memcpy_d2h.argtypes = [vp, vp, C.c_size_t]
        L.gc_memset.argtypes = [vp, C.c_int, C.c_size_t]
        L.gc_source_get.argtypes = [vp, u32, vp, vp, C.POINTER(C.c_int)]
        L.gc_cache_update.argtypes = [vp, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def check(rc):
    if rc != GC_OK:
        raise GcError("gcvol error %d: %s" % (rc, lib().gc_last_error().decode()))


def ptr(a):
    return a.ctypes.data_as(vp)


def device_count():
    n = C.c_int(0)
    check(lib().gc_device_count(C.byref(n)))
    return n.value


# ------------------------------------------------------------------ configs
def volume_cfg(dims, brick_size, covering, encoded_bytes, big_brick=(1, 1, 1)):
    vc = VolumeCfg()
    check(lib().gc_volume_cfg_init(C.byref(vc), ptr(np.asarray(dims, np.uint32)), u32(brick_size),
                                   u32(covering), ptr(np.asarray(big_brick, np.uint32)), u32(encoded_bytes)))
    return vc


def cache_cfg(cache_size, brick_size, max_requests=50, sync=True, empty_threshold=0.0):
    cc = CacheCfg()
    cc.cache_size[:] = list(cache_size)
    cc.block_size[:] = [brick_size] * 3
    cc.data_cache_flags = 1
    cc.max_requests = max_requests
    cc.sync_requests = 1 if sync else 0
    cc.empty_threshold = empty_threshold
    return cc


def pool_shape(budget, brick_bytes):
    out = np.zeros(3, np.uint32)
    check(lib().gc_pool_shape(budget, brick_bytes, ptr(out)))
    return tuple(int(v) for v in out)


def synth_params(seed, nb_blobs, sigma=(8.0, 40.0), amp=(60.0, 255.0), noise=8, threads=0):
    p = SynthParams()
    p.seed = seed
    p.nb_blobs = nb_blobs
    p.sigma_min, p.sigma_max = sigma
    p.amp_min, p.amp_max = amp
    p.noise = noise
    p.threads = threads
    return p


def synth_volume(params, dims, encoded_bytes):
    out = np.zeros((dims[2], dims[1], dims[0]), DT[encoded_bytes])
    check(lib().gc_synth_volume(C.byref(params), ptr(np.asarray(dims, np.uint32)), u32(encoded_bytes), ptr(out)))
    return out


def tf_seeded(seed, nb_knots=6):
    k = np.zeros((nb_knots, 5), np.float32)
    check(lib().gc_tf_seeded(seed, nb_knots, ptr(k)))
    return k


def tf_samples(knots, n):
    knots = np.ascontiguousarray(knots, np.float32)
    out = np.zeros((n, 4), np.float32)
    check(lib().gc_tf_samples(ptr(knots), u32(len(knots)), u32(n), ptr(out)))
    return out


def camera(rx, ry, distance):
    t = np.array([0.0, 0.0, -distance], np.float32)
    out = np.zeros(12, np.float32)
    check(lib().gc_camera_matrix(rx, ry, ptr(t), ptr(out)))
    return out


# ---------------------------------------------------------------- device mem
class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.p = vp()
        check(lib().gc_malloc(C.byref(self.p), self.nbytes))

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(lib().gc_memcpy_h2d(b.p, ptr(a), a.nbytes))
        return b

    def to_array(self, dtype, shape):
        out = np.zeros(shape, dtype)
        check(lib().gc_memcpy_d2h(ptr(out), self.p, out.nbytes))
        return out

    def free(self):
        if self.p:
            lib().gc_free(self.p)
            self.p = vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def mip_reduce(d_in, dims, encoded_bytes, d_out, stream=None):
    check(lib().gc_mip_reduce(d_in.p, ptr(np.asarray(dims, np.uint32)), u32(encoded_bytes), d_out.p, stream))


def synchronize():
    check(lib().gc_synchronize())


# ------------------------------------------------------------------ sources
class Source:
    def __init__(self, handle, keep=None):
        self.h = handle
        self._keep = keep

    @classmethod
    def pyramid(cls, vc, levels):
        levels = [np.ascontiguousarray(l) for l in levels]
        arr = (vp * len(levels))(*[l.ctypes.data for l in levels])
        h = vp()
        check(lib().gc_source_create_pyramid(C.byref(vc), arr, C.byref(h)))
        return cls(h, (levels, arr))

    @classmethod
    def synthetic(cls, vc, params):
        h = vp()
        check(lib().gc_source_create_synthetic(C.byref(vc), C.byref(params), C.byref(h)))
        return cls(h)

    def get(self, vc, level, bb_coord):
        bb = [vc.big_brick_size[a] for a in range(3)]
        bs = vc.brick_size[0]
        out = np.zeros((bs * bb[2], bs * bb[1], bs * bb[0]), DT[vc.encoded_bytes])
        st = C.c_int(0)
        check(lib().gc_source_get(self.h, level, ptr(np.asarray(bb_coord, np.uint32)), ptr(out), C.byref(st)))
        return st.value, out

    def close(self):
        if self.h:
            lib().gc_source_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -------------------------------------------------------------------- cache
class Cache:
    def __init__(self, vc, cc, source, device=0):
        self.vc, self.cc, self.source = vc, cc, source
        self.h = vp()
        check(lib().gc_cache_create(C.byref(vc), C.byref(cc), source.h, device, C.byref(self.h)))
        self.g = Geometry()
        check(lib().gc_cache_geometry(self.h, C.byref(self.g)))
        self.cells = int(self.g.mrpd_cells)
        self.nslots = int(self.g.nslots)

    def close(self):
        if self.h:
            lib().gc_cache_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ts(self):
        t = u32(0)
        check(lib().gc_cache_timestamp(self.h, C.byref(t)))
        return int(t.value)

    def stream(self):
        s = vp()
        check(lib().gc_cache_stream(self.h, C.byref(s)))
        return s

    def update(self):
        d = C.c_int(0)
        check(lib().gc_cache_update(self.h, C.byref(d)))
        return bool(d.value)

    def completude(self):
        f = f32(0)
        check(lib().gc_cache_completude(self.h, C.byref(f)))
        return float(f.value)

    def load_all(self, d_levels):
        arr = (vp * len(d_levels))(*[d.p.value for d in d_levels])
        check(lib().gc_cache_load_all(self.h, arr))

    def mrpd(self):
        out = np.zeros((self.cells, 4), np.uint32)
        check(lib().gc_cache_read_mrpd(self.h, ptr(out)))
        return out

    def request(self):
        out = np.zeros(self.cells, np.uint32)
        check(lib().gc_cache_read_request(self.h, ptr(out)))
        return out

    def write_request(self, a):
        a = np.ascontiguousarray(a, np.uint32)
        check(lib().gc_cache_write_request(self.h, ptr(a)))

    def usage(self):
        out = np.zeros(self.nslots, np.uint32)
        check(lib().gc_cache_read_usage(self.h, ptr(out)))
        return out

    def write_usage(self, a):
        a = np.ascontiguousarray(a, np.uint32)
        check(lib().gc_cache_write_usage(self.h, ptr(a)))

    def lru(self):
        out = np.zeros((self.nslots, 7), np.uint32)
        check(lib().gc_cache_read_lru(self.h, ptr(out)))
        return out

    def pool(self):
        bs = self.g.brick_size
        d = [int(self.g.cache_size[a]) * bs for a in range(3)]
        out = np.zeros((d[2], d[1], d[0]), DT[self.vc.encoded_bytes])
        check(lib().gc_cache_read_pool(self.h, ptr(out)))
        return out

    def asked(self):
        n, nne = u32(0), u32(0)
        check(lib().gc_cache_read_asked(self.h, None, C.byref(n), C.byref(nne)))
        out = np.zeros((max(n.value, 1), 4), np.uint32)
        check(lib().gc_cache_read_asked(self.h, ptr(out), C.byref(n), C.byref(nne)))
        return out[: n.value].copy(), int(nne.value)


class Frame:
    def __init__(self, width, height, precision=32):
        self.w, self.h_, self.precision = width, height, precision
        self.h = vp()
        check(lib().gc_frame_create(width, height, precision, C.byref(self.h)))

    def reset(self, stream=None):
        check(lib().gc_frame_reset(self.h, stream))

    def pixels(self):
        out = np.zeros((self.h_, self.w), np.uint32)
        check(lib().gc_frame_read_pixels(self.h, ptr(out)))
        return out

    def state(self):
        t = np.zeros((self.h_, self.w), np.float32)
        c = np.zeros((self.h_, self.w, 4), np.float64 if self.precision == 64 else np.float32)
        check(lib().gc_frame_read_state(self.h, ptr(t), ptr(c)))
        return t, c

    def close(self):
        if self.h:
            lib().gc_frame_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_view(width, height, inv_mv, alpha_stop=0.95, precision=32, tile=0,
              bbox=((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))):
    v = View()
    v.width, v.height = width, height
    v.inv_mv[:] = [float(x) for x in np.asarray(inv_mv, np.float32).reshape(12)]
    v.bbox_min[:] = list(bbox[0])
    v.bbox_max[:] = list(bbox[1])
    v.alpha_stop = alpha_stop
    v.precision = precision
    v.tile = tile
    return v


def render(cache, view, d_tf, tf_size, frame, d_stats=None, stream=None):
    check(lib().gc_render_dvr(cache.h, C.byref(view), d_tf.p, tf_size, frame.h,
                              d_stats.p if d_stats is not None else None, stream))


def unpack_rgba(px):
    px = np.asarray(px, np.uint32)
    return np.stack([(px >> s) & 0xFF for s in (0, 8, 16, 24)], axis=-1).astype(np.int32)
