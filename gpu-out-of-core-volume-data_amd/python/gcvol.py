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
MAXT = 3

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
        ("nb_tables", u32), ("table_cache_size", (u32 * 3) * MAXT), ("table_block_size", (u32 * 3) * MAXT),
    ]


class SynthParams(C.Structure):
    _fields_ = [
        ("seed", u64), ("nb_blobs", u32), ("sigma_min", f32), ("sigma_max", f32),
        ("amp_min", f32), ("amp_max", f32), ("noise", u32), ("threads", u32),
    ]


class TableGeometry(C.Structure):
    _fields_ = [
        ("block", u32 * 3), ("blocks", u32 * 3), ("surf", u32 * 3), ("nblocks", u32),
        ("virt", (u32 * 3) * MAXL), ("real", (u32 * 3) * MAXL),
    ]


class Geometry(C.Structure):
    _fields_ = [
        ("nb_levels", u32), ("brick_size", u32), ("covering", u32), ("inner", u32), ("encoded_bytes", u32),
        ("level_dim", (u32 * 3) * MAXL), ("level_coord", u32 * MAXL), ("virt_vox", (u32 * 3) * MAXL),
        ("ior", (f32 * 3) * MAXL), ("rdv", (f32 * 3) * MAXL), ("one_over", f32 * 3),
        ("mrpd_dim", u32 * 3), ("cache_size", u32 * 3), ("nslots", u32), ("pad", u32), ("mrpd_cells", u64),
        ("nb_tables", u32), ("top_dim", (u32 * 3) * MAXL), ("top_coord", u32 * MAXL), ("top_cells_dim", u32 * 3),
        ("top_cells", u64), ("table", TableGeometry * MAXT),
    ]


class View(C.Structure):
    _fields_ = [
        ("width", u32), ("height", u32), ("inv_mv", f32 * 12), ("bbox_min", f32 * 3), ("bbox_max", f32 * 3),
        ("alpha_stop", f32), ("precision", u32), ("tile", u32), ("refill_min", u32), ("march", u32),
        ("resume", u32), ("compositing", u32),
    ]


MSG_TF_MAX_ENTRIES = 2047


class ViewRequest(C.Structure):
    _fields_ = [("inv_mv", f32 * 12), ("bbox_min", f32 * 3), ("bbox_max", f32 * 3), ("tf_entries", u32),
                ("have", u32), ("tf", f32 * (MSG_TF_MAX_ENTRIES * 4))]


class RenderStats(C.Structure):
    _fields_ = [("mapped_samples", u64), ("brick_visits", u64), ("pixels", u64), ("unmapped_rays", u64)]


class LaunchRecord(C.Structure):
    _fields_ = [("launch", u32), ("kernel", u32), ("list_rays", u32), ("group", u32), ("stats", RenderStats),
                ("kernel_ms", C.c_double), ("bricks_loaded", u32), ("handed_over", u32)]


class FrameReport(C.Structure):
    _fields_ = [
        ("launches", u32), ("updates", u32), ("complete", u32), ("reserved", u32), ("bricks_loaded", u64),
        ("stats", RenderStats), ("kernel_ms", C.c_double), ("update_ms", C.c_double),
        ("lane_launches", u32), ("reserved2", u32), ("lane_stats", RenderStats), ("lane_kernel_ms", C.c_double),
    ]

    def as_dict(self):
        return {
            "launches": self.launches, "updates": self.updates, "complete": bool(self.complete),
            "bricks_loaded": int(self.bricks_loaded), "mapped_samples": int(self.stats.mapped_samples),
            "brick_visits": int(self.stats.brick_visits), "pixels": int(self.stats.pixels),
            "unmapped_rays": int(self.stats.unmapped_rays), "kernel_ms": self.kernel_ms,
            "update_ms": self.update_ms,
            "lane_launches": self.lane_launches, "lane_kernel_ms": self.lane_kernel_ms,
            "lane_mapped_samples": int(self.lane_stats.mapped_samples),
            "lane_brick_visits": int(self.lane_stats.brick_visits), "lane_pixels": int(self.lane_stats.pixels),
        }


class CacheView(C.Structure):
    _fields_ = [
        ("mrpd", vp), ("request", vp), ("usage", vp), ("pool", vp), ("geometry", vp),
        ("timestamp", u32), ("brick_size", u32), ("covering", u32), ("encoded_bytes", u32),
        ("nb_tables", u32), ("table_words", vp * MAXT), ("table_usage", vp * MAXT),
    ]


class ServerCfg(C.Structure):
    _fields_ = [("pool_budget", u64), ("max_views", u32), ("slot_pixels", u32), ("precision", u32), ("tile", u32),
                ("march", u32), ("plain", u32), ("graph", u32), ("fixed_views", u32)]


class ServerResult(C.Structure):
    _fields_ = [("id", u64), ("volume", u32), ("width", u32), ("height", u32), ("launches", u32),
                ("submit_ms", C.c_double), ("latency_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", u64), ("completed", u64), ("launches", u64), ("samples", u64),
                ("pools_created", u64), ("pools_evicted", u64), ("bricks_loaded", u64), ("pressure_events", u64),
                ("iterations", u64), ("idle_iterations", u64), ("launch_ms", C.c_double), ("update_ms", C.c_double),
                ("fetch_ms", C.c_double), ("ray_starts", u64)]


class CacheCounters(C.Structure):
    _fields_ = [("updates", u64), ("loads", u64), ("bricks_loaded", u64), ("requests", u64), ("slots", u32),
                ("last_used", u32), ("last_requested", u32), ("reserved", u32), ("fetch_ms", C.c_double)]


EXPORTS = [
    "gc_last_error", "gc_version", "gc_volume_cfg_init", "gc_pool_shape", "gc_cfg_get", "gc_volume_cfg_load",
    "gc_device_count", "gc_set_device", "gc_malloc", "gc_free", "gc_memcpy_h2d", "gc_memcpy_d2h", "gc_memcpy_d2d",
    "gc_memset", "gc_synchronize", "gc_stream_create", "gc_stream_destroy", "gc_stream_synchronize",
    "gc_mip_reduce", "gc_brick_cut", "gc_source_create_pyramid", "gc_source_create_synthetic",
    "gc_source_create_callback", "gc_source_destroy", "gc_source_get", "gc_synth_volume",
    "gc_cache_create", "gc_cache_destroy", "gc_cache_update", "gc_cache_completude", "gc_cache_load_all",
    "gc_cache_timestamp", "gc_cache_stream", "gc_cache_view_get", "gc_cache_geometry", "gc_geometry_compute",
    "gc_cache_read_mrpd", "gc_cache_read_cells", "gc_cache_read_table", "gc_cache_read_table_usage",
    "gc_cache_write_table_usage", "gc_cache_read_table_lru", "gc_cache_unreferenced",
    "gc_cache_read_request", "gc_cache_write_request", "gc_cache_read_usage", "gc_cache_write_usage",
    "gc_cache_read_lru", "gc_cache_read_pool", "gc_cache_read_asked", "gc_frame_create", "gc_frame_destroy",
    "gc_frame_reset", "gc_frame_read_pixels", "gc_frame_read_state", "gc_frame_buffers", "gc_render_dvr",
    "gc_tf_samples", "gc_tf_seeded", "gc_camera_matrix", "gc_cache_wait_idle", "gc_render_frame",
    "gc_batch_create", "gc_batch_destroy", "gc_batch_set_graph", "gc_batch_render", "gc_batch_read_pixels",
    "gc_batch_set_launch", "gc_batch_begin", "gc_batch_set_view", "gc_batch_clear_view", "gc_batch_launch",
    "gc_batch_pending", "gc_batch_stats", "gc_batch_read_state", "gc_cache_set_max_requests",
    "gc_cache_set_ray_list_max", "gc_cache_set_ray_group", "gc_cache_set_ray_handoff", "gc_cache_get_ray_handoff",
    "gc_cache_set_launch_log",
    "gc_cache_busy", "gc_source_create_gpu_mip", "gc_server_create", "gc_server_destroy", "gc_server_add_volume", "gc_server_submit",
    "gc_server_poll", "gc_server_take_pixels", "gc_server_drain", "gc_server_stats_get", "gc_cache_counters_get",
    "gc_source_create_synthetic_gpu", "gc_source_get_device", "gc_render_views_noalloc", "gc_cache_probe",
    "gc_brick_key", "gc_brick_key_unpack", "gc_store_brick_path", "gc_lz4_bound", "gc_lz4_compress",
    "gc_lz4_decompress", "gc_store_write", "gc_source_create_store", "gc_store_read_meta", "gc_store_stats",
    "gc_msg_encode_camera", "gc_msg_encode_transfer_function", "gc_msg_encode_bounding_box", "gc_msg_id",
    "gc_view_request_init", "gc_view_request_apply", "gc_view_request_view", "gc_source_create_gpu_pyramid",
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
        L.gc_brick_key.restype = u64
        L.gc_brick_key_unpack.restype = None
        L.gc_lz4_bound.restype = u64
        # Every export gets explicit argument types: without them ctypes passes
        # Python ints as 32-bit C ints and silently truncates pointers.
        P, I, SZ = vp, C.c_int, C.c_size_t
        sig = {
            "gc_last_error": [], "gc_version": [],
            "gc_volume_cfg_init": [C.POINTER(VolumeCfg), P, u32, u32, P, u32],
            "gc_pool_shape": [u64, u64, P], "gc_cfg_get": [C.c_char_p, C.c_char_p, C.c_char_p, SZ],
            "gc_volume_cfg_load": [C.c_char_p, P, C.c_char_p, SZ],
            "gc_device_count": [C.POINTER(I)], "gc_set_device": [I],
            "gc_malloc": [C.POINTER(vp), SZ], "gc_free": [P], "gc_memcpy_h2d": [P, P, SZ],
            "gc_memcpy_d2h": [P, P, SZ], "gc_memcpy_d2d": [P, P, SZ], "gc_memset": [P, I, SZ],
            "gc_synchronize": [],
            "gc_stream_create": [C.POINTER(vp)], "gc_stream_destroy": [P], "gc_stream_synchronize": [P],
            "gc_mip_reduce": [P, P, u32, P, P], "gc_brick_cut": [P, P, u32, u32, u32, P, u32, P, P],
            "gc_source_create_pyramid": [C.POINTER(VolumeCfg), P, C.POINTER(vp)],
            "gc_source_create_synthetic": [C.POINTER(VolumeCfg), C.POINTER(SynthParams), C.POINTER(vp)],
            "gc_source_create_callback": [C.POINTER(VolumeCfg), P, P, C.POINTER(vp)],
            "gc_source_destroy": [P], "gc_source_get": [P, u32, P, P, C.POINTER(I)],
            "gc_synth_volume": [C.POINTER(SynthParams), P, u32, P],
            "gc_cache_create": [C.POINTER(VolumeCfg), C.POINTER(CacheCfg), P, I, C.POINTER(vp)],
            "gc_cache_destroy": [P], "gc_cache_update": [P, C.POINTER(I)],
            "gc_cache_completude": [P, C.POINTER(f32)], "gc_cache_load_all": [P, P],
            "gc_cache_timestamp": [P, C.POINTER(u32)], "gc_cache_stream": [P, C.POINTER(vp)],
            "gc_cache_counters_get": [P, C.POINTER(CacheCounters)],
            "gc_cache_view_get": [P, C.POINTER(CacheView)], "gc_cache_geometry": [P, C.POINTER(Geometry)],
            "gc_geometry_compute": [C.POINTER(VolumeCfg), C.POINTER(CacheCfg), C.POINTER(Geometry)],
            "gc_cache_read_mrpd": [P, P], "gc_cache_read_cells": [P, P], "gc_cache_read_table": [P, u32, P],
            "gc_cache_read_table_usage": [P, u32, P], "gc_cache_write_table_usage": [P, u32, P],
            "gc_cache_read_table_lru": [P, u32, P], "gc_cache_unreferenced": [P, P], "gc_cache_read_request": [P, P], "gc_cache_write_request": [P, P],
            "gc_cache_read_usage": [P, P], "gc_cache_write_usage": [P, P], "gc_cache_read_lru": [P, P],
            "gc_cache_read_pool": [P, P], "gc_cache_read_asked": [P, P, C.POINTER(u32), C.POINTER(u32)],
            "gc_frame_create": [u32, u32, u32, C.POINTER(vp)], "gc_frame_destroy": [P],
            "gc_frame_reset": [P, P], "gc_frame_read_pixels": [P, P], "gc_frame_read_state": [P, P, P],
            "gc_frame_buffers": [P, P, P, P],
            "gc_render_dvr": [P, C.POINTER(View), P, u32, P, P, P],
            "gc_tf_samples": [P, u32, u32, P], "gc_tf_seeded": [u64, u32, P],
            "gc_camera_matrix": [f32, f32, P, P], "gc_cache_wait_idle": [P],
            "gc_render_frame": [P, C.POINTER(View), P, u32, P, u32, C.POINTER(FrameReport)],
            "gc_batch_create": [P, u32, u64, u32, C.POINTER(vp)], "gc_batch_destroy": [P],
            "gc_batch_set_graph": [P, C.c_int],
            "gc_batch_render": [P, C.POINTER(View), C.POINTER(vp), P, u32, u32, C.POINTER(FrameReport)],
            "gc_batch_read_pixels": [P, u32, P], "gc_cache_set_max_requests": [P, u32],
            "gc_cache_set_ray_list_max": [P, u32], "gc_cache_set_ray_group": [P, u32],
            "gc_cache_set_ray_handoff": [P, u32, u32, u32],
            "gc_cache_get_ray_handoff": [P, P, P, P],
            "gc_cache_set_launch_log": [P, P, u32, P],
            "gc_batch_set_launch": [P, u32, u32, u32],
            "gc_batch_begin": [P, C.POINTER(View), C.POINTER(vp), P, u32],
            "gc_batch_set_view": [P, u32, C.POINTER(View), P, u32], "gc_batch_clear_view": [P, u32],
            "gc_batch_launch": [P, u32, C.POINTER(C.c_double)], "gc_batch_pending": [P, P],
            "gc_batch_stats": [P, C.POINTER(RenderStats)], "gc_batch_read_state": [P, u32, P, P],
            "gc_cache_busy": [P, C.POINTER(C.c_int)],
            "gc_source_create_gpu_mip": [C.POINTER(VolumeCfg), P, C.c_int, u64, C.POINTER(vp)],
            "gc_source_create_synthetic_gpu": [C.POINTER(VolumeCfg), C.POINTER(SynthParams), C.c_int, C.POINTER(vp)],
            "gc_source_get_device": [P, P, u32, P, C.c_float, P],
            "gc_server_create": [C.POINTER(ServerCfg), C.c_int, C.POINTER(vp)], "gc_server_destroy": [P],
            "gc_server_add_volume": [P, C.POINTER(VolumeCfg), P, u64, u32, C.POINTER(u32)],
            "gc_server_submit": [P, u32, C.POINTER(View), P, u32, C.POINTER(u64)],
            "gc_server_poll": [P, C.POINTER(ServerResult), u32, C.POINTER(u32)],
            "gc_server_take_pixels": [P, u64, P], "gc_server_drain": [P],
            "gc_server_stats_get": [P, C.POINTER(ServerStats)],
            "gc_render_views_noalloc": [P, C.POINTER(View), C.POINTER(vp), P, u32, C.POINTER(vp), C.POINTER(vp),
                                        C.POINTER(vp), P, P],
            "gc_cache_probe": [P, u32, P, P, u32, P, P],
            "gc_brick_key": [u32, u32, u32, u32], "gc_brick_key_unpack": [u64, C.POINTER(u32), P],
            "gc_store_brick_path": [C.c_char_p, u32, u32, P, C.c_char_p, SZ],
            "gc_lz4_bound": [u64], "gc_lz4_compress": [P, u64, P, u64, C.POINTER(u64)],
            "gc_lz4_decompress": [P, u64, P, u64],
            "gc_store_write": [C.POINTER(VolumeCfg), P, C.c_char_p, C.c_int, C.c_float, P, u32, u32],
            "gc_source_create_store": [C.POINTER(VolumeCfg), C.c_char_p, u32, C.POINTER(vp)],
            "gc_store_read_meta": [C.c_char_p, u32, P, C.POINTER(u32), P, C.POINTER(u32)],
            "gc_store_stats": [P, P],
            "gc_msg_encode_camera": [P, P, SZ, C.POINTER(SZ)],
            "gc_msg_encode_transfer_function": [P, u32, P, SZ, C.POINTER(SZ)],
            "gc_msg_encode_bounding_box": [P, P, P, SZ, C.POINTER(SZ)],
            "gc_msg_id": [P, SZ, C.POINTER(u32)],
            "gc_view_request_init": [C.POINTER(ViewRequest)], "gc_view_request_apply": [C.POINTER(ViewRequest), P, SZ],
            "gc_view_request_view": [C.POINTER(ViewRequest), u32, u32, u32, C.POINTER(View)],
            "gc_source_create_gpu_pyramid": [C.POINTER(VolumeCfg), C.POINTER(SynthParams), C.c_int, u32,
                                             C.POINTER(vp), C.POINTER(C.c_double)],
        }
        missing = [n for n in EXPORTS if n not in sig]
        if missing:
            raise GcError("binding lacks argument types for %s" % missing)
        for name, args in sig.items():
            getattr(L, name).argtypes = args
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


def volume_cfg_load(cfg_path):
    """load_volume_configuration of a GcCore .cfg: (gc_volume_cfg, volume path)."""
    vc = VolumeCfg()
    buf = C.create_string_buffer(4096)
    check(lib().gc_volume_cfg_load(str(cfg_path).encode(), C.byref(vc), buf, 4096))
    return vc, buf.value.decode()


def cache_cfg(cache_size, brick_size, max_requests=50, sync=True, empty_threshold=0.0, tables=()):
    """tables: page-table caches as (CacheSize[k+1], BlockSize[k+1]) pairs in
    CacheConfiguration order (k = 0 next to the brick pool)."""
    cc = CacheCfg()
    cc.cache_size[:] = list(cache_size)
    cc.block_size[:] = [brick_size] * 3
    cc.data_cache_flags = 1
    cc.max_requests = max_requests
    cc.sync_requests = 1 if sync else 0
    cc.empty_threshold = empty_threshold
    cc.nb_tables = len(tables)
    for k, (cs, bl) in enumerate(tables):
        cc.table_cache_size[k][:] = [int(v) for v in cs]
        cc.table_block_size[k][:] = [int(v) for v in bl]
    return cc


def geometry_compute(vc, cc):
    g = Geometry()
    check(lib().gc_geometry_compute(C.byref(vc), C.byref(cc), C.byref(g)))
    return g


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
    def synthetic(cls, vc, params, device=None):
        """Seeded synthetic volume; with `device`, bricks are synthesised on
        that GPU straight into the cache's staging buffer."""
        h = vp()
        if device is None:
            check(lib().gc_source_create_synthetic(C.byref(vc), C.byref(params), C.byref(h)))
        else:
            check(lib().gc_source_create_synthetic_gpu(C.byref(vc), C.byref(params), int(device), C.byref(h)))
        return cls(h)

    def get_device(self, vc, keys, threshold=0.0):
        """Bricks (bx, by, bz, level) of a device source, synthesised on the
        GPU: (n, bs, bs, bs) voxels and the per-brick empty flags."""
        keys = np.ascontiguousarray(keys, np.uint32).reshape(-1, 4)
        n, bs = len(keys), vc.brick_size[0]
        d = DeviceBuffer(max(1, n * bs ** 3 * vc.encoded_bytes))
        empty = np.zeros(max(1, n), np.uint8)
        check(lib().gc_source_get_device(self.h, ptr(keys), n, d.p, float(threshold), ptr(empty)))
        out = d.to_array(DT[vc.encoded_bytes], (n, bs, bs, bs))
        d.free()
        return out, empty[:n]

    @classmethod
    def gpu_pyramid(cls, vc, params, device=0, slab_planes=0):
        """Synthetic level 0 with its mip pyramid built on the GPU (cold
        build); returns (source, build milliseconds)."""
        h = vp()
        ms = C.c_double(0)
        check(lib().gc_source_create_gpu_pyramid(C.byref(vc), C.byref(params), int(device), slab_planes,
                                                 C.byref(h), C.byref(ms)))
        return cls(h), ms.value

    @classmethod
    def gpu_mip(cls, vc, level0, device=0, memo_bytes=256 << 20):
        """Mip on miss over a level-0 source (kept alive with this source)."""
        h = vp()
        check(lib().gc_source_create_gpu_mip(C.byref(vc), level0.h, device, memo_bytes, C.byref(h)))
        return cls(h, keep=level0)

    @classmethod
    def store(cls, vc, volume_dir, cache_bricks=0):
        """The on-disk brick store under volume_dir (bricks_B_B_B/...), read
        through a host LRU of cache_bricks big bricks (0: 32768)."""
        h = vp()
        check(lib().gc_source_create_store(C.byref(vc), os.fsencode(volume_dir), cache_bricks, C.byref(h)))
        return cls(h)

    def store_stats(self):
        """(host-cache hits, file loads, big bricks held, Empty answers)."""
        out = np.zeros(4, np.uint64)
        check(lib().gc_store_stats(self.h, ptr(out)))
        return tuple(int(x) for x in out)

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
        self.top_cells = int(self.g.top_cells)
        self.nb_tables = int(self.g.nb_tables)
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

    def counters(self):
        c = CacheCounters()
        check(lib().gc_cache_counters_get(self.h, C.byref(c)))
        return {f: getattr(c, f) for f, _ in CacheCounters._fields_ if f != "reserved"}

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
        out = np.zeros((self.top_cells, 4), np.uint32)
        check(lib().gc_cache_read_mrpd(self.h, ptr(out)))
        return out

    def cell_table(self):
        out = np.zeros(self.top_cells, np.uint32)
        check(lib().gc_cache_read_cells(self.h, ptr(out)))
        return out

    def table(self, i):
        """Entries of page-table cache i (i = 0 next to the MRPD) in the
        reference's surface layout: (entries, 4)."""
        t = self.g.table[i]
        out = np.zeros((int(t.surf[0]) * int(t.surf[1]) * int(t.surf[2]), 4), np.uint32)
        check(lib().gc_cache_read_table(self.h, i, ptr(out)))
        return out

    def table_usage(self, i):
        out = np.zeros(int(self.g.table[i].nblocks), np.uint32)
        check(lib().gc_cache_read_table_usage(self.h, i, ptr(out)))
        return out

    def write_table_usage(self, i, a):
        a = np.ascontiguousarray(a, np.uint32)
        check(lib().gc_cache_write_table_usage(self.h, i, ptr(a)))

    def table_lru(self, i):
        out = np.zeros((int(self.g.table[i].nblocks), 7), np.uint32)
        check(lib().gc_cache_read_table_lru(self.h, i, ptr(out)))
        return out

    def unreferenced(self):
        n = u32(0)
        check(lib().gc_cache_unreferenced(self.h, C.byref(n)))
        return int(n.value)

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

    def probe(self, mode, pos, levels, init=0.0):
        """Device API probes (mode 0 gc_get_normalized, 1 gc_get):
        (values, statuses)."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        n = len(pos)
        d_pos = DeviceBuffer.from_array(pos)
        d_lv = DeviceBuffer.from_array(np.ascontiguousarray(levels, np.uint32))
        d_val = DeviceBuffer.from_array(np.full(n, init, np.float32))
        d_st = DeviceBuffer(4 * n)
        check(lib().gc_cache_probe(self.h, mode, d_pos.p, d_lv.p, n, d_val.p, d_st.p))
        return d_val.to_array(np.float32, (n,)), d_st.to_array(np.uint32, (n,))

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
              bbox=((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)), refill_min=0, march=0, resume=0, compositing=0):
    v = View()
    v.width, v.height = width, height
    v.inv_mv[:] = [float(x) for x in np.asarray(inv_mv, np.float32).reshape(12)]
    v.bbox_min[:] = list(bbox[0])
    v.bbox_max[:] = list(bbox[1])
    v.alpha_stop = alpha_stop
    v.precision = precision
    v.tile = tile
    v.refill_min = refill_min
    v.march = march
    v.resume = resume
    v.compositing = compositing
    return v


def render(cache, view, d_tf, tf_size, frame, d_stats=None, stream=None):
    check(lib().gc_render_dvr(cache.h, C.byref(view), d_tf.p, tf_size, frame.h,
                              d_stats.p if d_stats is not None else None, stream))


def render_frame(cache, view, d_tf, tf_size, frame, max_launches=100000):
    rep = FrameReport()
    check(lib().gc_render_frame(cache.h, C.byref(view), d_tf.p, tf_size, frame.h, max_launches, C.byref(rep)))
    return rep


class Batch:
    """View slots of one cache rendered together (gc_batch_*)."""

    def __init__(self, cache, max_views, max_pixels, precision=32):
        self.h = vp()
        self.cache = cache
        self.max_views = max_views
        self.precision = precision
        check(lib().gc_batch_create(cache.h, max_views, max_pixels, precision, C.byref(self.h)))
        self.sizes = {}

    def set_graph(self, enable):
        check(lib().gc_batch_set_graph(self.h, 1 if enable else 0))

    def set_launch(self, tile=0, march=0, refill_min=0):
        check(lib().gc_batch_set_launch(self.h, tile, march, refill_min))

    @staticmethod
    def _arrays(views, d_tfs, tf_sizes):
        n = len(views)
        return n, (View * n)(*views), (vp * n)(*[t.p for t in d_tfs]), np.asarray(tf_sizes, np.uint32)

    def begin(self, views, d_tfs, tf_sizes):
        n, va, ta, sa = self._arrays(views, d_tfs, tf_sizes)
        check(lib().gc_batch_begin(self.h, va, ta, ptr(sa), n))
        self.sizes = {i: (v.width, v.height) for i, v in enumerate(views)}

    def render(self, views, d_tfs, tf_sizes, max_launches=100000):
        n, va, ta, sa = self._arrays(views, d_tfs, tf_sizes)
        rep = FrameReport()
        check(lib().gc_batch_render(self.h, va, ta, ptr(sa), n, max_launches, C.byref(rep)))
        self.sizes = {i: (v.width, v.height) for i, v in enumerate(views)}
        return rep

    def set_view(self, slot, view, d_tf, tf_size):
        check(lib().gc_batch_set_view(self.h, slot, C.byref(view), d_tf.p, tf_size))
        self.sizes[slot] = (view.width, view.height)

    def clear_view(self, slot):
        check(lib().gc_batch_clear_view(self.h, slot))
        self.sizes.pop(slot, None)

    def launch(self, starting_only=False):
        ms = C.c_double(0)
        check(lib().gc_batch_launch(self.h, 1 if starting_only else 0, C.byref(ms)))
        return ms.value

    def pending(self):
        out = np.zeros(self.max_views, np.uint32)
        check(lib().gc_batch_pending(self.h, ptr(out)))
        return out

    def stats(self):
        st = RenderStats()
        check(lib().gc_batch_stats(self.h, C.byref(st)))
        return st

    def pixels(self, i):
        w, h = self.sizes[i]
        out = np.zeros(w * h, np.uint32)
        check(lib().gc_batch_read_pixels(self.h, i, ptr(out)))
        return out

    def state(self, i):
        w, h = self.sizes[i]
        tb = np.zeros(w * h, np.float32)
        cs = np.zeros(w * h * 4, np.float64 if self.precision == 64 else np.float32)
        check(lib().gc_batch_read_state(self.h, i, ptr(tb), ptr(cs)))
        return tb, cs

    def close(self):
        if self.h:
            lib().gc_batch_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving loop (gc_server_*)."""

    def __init__(self, pool_budget, max_views=64, slot_pixels=2048 * 2048, precision=32, tile=0, march=0,
                 plain=False, graph=False, device=0, fixed_views=False):
        cfg = ServerCfg(pool_budget, max_views, slot_pixels, precision, tile, march, 1 if plain else 0,
                        1 if graph else 0, 1 if fixed_views else 0)
        self.h = vp()
        self.keep = []
        check(lib().gc_server_create(C.byref(cfg), device, C.byref(self.h)))

    def add_volume(self, vc, source, pool_bytes, max_requests=1024):
        v = u32(0)
        check(lib().gc_server_add_volume(self.h, C.byref(vc), source.h, pool_bytes, max_requests, C.byref(v)))
        self.keep.append(source)
        return v.value

    def submit(self, volume, view, tf):
        tf = np.ascontiguousarray(tf, np.float32)
        i = u64(0)
        check(lib().gc_server_submit(self.h, volume, C.byref(view), ptr(tf), len(tf), C.byref(i)))
        return i.value

    def poll(self, max_results=256):
        out = (ServerResult * max_results)()
        n = u32(0)
        check(lib().gc_server_poll(self.h, out, max_results, C.byref(n)))
        return [out[k] for k in range(n.value)]

    def take_pixels(self, rid, width, height):
        px = np.zeros(width * height, np.uint32)
        check(lib().gc_server_take_pixels(self.h, rid, ptr(px)))
        return px

    def drain(self):
        check(lib().gc_server_drain(self.h))

    def stats(self):
        st = ServerStats()
        check(lib().gc_server_stats_get(self.h, C.byref(st)))
        return st

    def close(self):
        if self.h:
            lib().gc_server_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_views_noalloc(cache, views, d_tfs, tf_sizes, d_pixels, d_tbufs, d_cstates, d_stats=None, stream=None):
    """One launch over the views into caller device buffers (DeviceBuffer
    lists), no device allocation (gc_render_views_noalloc)."""
    n = len(views)
    arr = lambda bufs: (vp * n)(*[b.p for b in bufs])
    check(lib().gc_render_views_noalloc(cache.h, (View * n)(*views), arr(d_tfs), ptr(np.asarray(tf_sizes, np.uint32)),
                                        n, arr(d_pixels), arr(d_tbufs), arr(d_cstates),
                                        d_stats.p if d_stats is not None else None, stream))


# ----------------------------------------------------------- brick store
def brick_key(level, x, y, z):
    """Bkey: (level, x, y, z) as four 16-bit fields, level lowest."""
    return int(lib().gc_brick_key(level, x, y, z))


def brick_key_unpack(key):
    lv = u32(0)
    xyz = np.zeros(3, np.uint32)
    lib().gc_brick_key_unpack(int(key), C.byref(lv), ptr(xyz))
    return int(lv.value), int(xyz[0]), int(xyz[1]), int(xyz[2])


def store_brick_path(volume_dir, brick_size, level, xyz):
    buf = C.create_string_buffer(4096)
    check(lib().gc_store_brick_path(os.fsencode(volume_dir), brick_size, level,
                                    ptr(np.asarray(xyz, np.uint32)), buf, 4096))
    return os.fsdecode(buf.value)


def lz4_compress(data):
    src = np.frombuffer(bytes(data), np.uint8)
    cap = int(lib().gc_lz4_bound(len(src)))
    dst = np.zeros(max(cap, 1), np.uint8)
    n = u64(0)
    check(lib().gc_lz4_compress(ptr(src) if len(src) else None, len(src), ptr(dst), cap, C.byref(n)))
    return dst[:n.value].tobytes()


def lz4_decompress(data, size):
    src = np.frombuffer(bytes(data), np.uint8)
    dst = np.zeros(max(size, 1), np.uint8)
    check(lib().gc_lz4_decompress(ptr(src), len(src), ptr(dst), size))
    return dst[:size].tobytes()


def store_write(vc, source, volume_dir, compression=True, empty_threshold=0.0, histogram=None, threads=0):
    """Write the source's bricks and bricks.xml as the reference's store."""
    h = None if histogram is None else np.ascontiguousarray(histogram, np.float32)
    check(lib().gc_store_write(C.byref(vc), source.h, os.fsencode(volume_dir), 1 if compression else 0,
                               float(empty_threshold), None if h is None else ptr(h), 0 if h is None else len(h),
                               threads))


def store_read_meta(volume_dir, brick_size):
    """(empty keys, histogram) of a store's bricks.xml."""
    ne, nh = u32(0), u32(0)
    check(lib().gc_store_read_meta(os.fsencode(volume_dir), brick_size, None, C.byref(ne), None, C.byref(nh)))
    keys = np.zeros(max(1, ne.value), np.uint64)
    histo = np.zeros(max(1, nh.value), np.float32)
    check(lib().gc_store_read_meta(os.fsencode(volume_dir), brick_size, ptr(keys), C.byref(ne), ptr(histo),
                                   C.byref(nh)))
    return keys[:ne.value], histo[:nh.value]


# ------------------------------------------------- view-request messages
MSG_CONFIGURATION, MSG_H264_CHUNKS, MSG_CAMERA, MSG_TRANSFER_FUNCTION, MSG_BOUNDING_BOX = 1, 2, 3, 4, 5


def _msg(fn, *args):
    out = C.create_string_buffer(32768)
    n = C.c_size_t(0)
    check(fn(*args, out, 32768, C.byref(n)))
    return out.raw[:n.value]


def msg_camera(inv_view):
    return _msg(lib().gc_msg_encode_camera, ptr(np.ascontiguousarray(inv_view, np.float32).reshape(12)))


def msg_transfer_function(rgba):
    rgba = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4)
    return _msg(lib().gc_msg_encode_transfer_function, ptr(rgba), len(rgba))


def msg_bounding_box(bmin, bmax):
    return _msg(lib().gc_msg_encode_bounding_box, ptr(np.asarray(bmin, np.float32)), ptr(np.asarray(bmax, np.float32)))


class Request:
    """A client's view request folded from the reference's messages."""

    def __init__(self):
        self.r = ViewRequest()
        check(lib().gc_view_request_init(C.byref(self.r)))

    def apply(self, msg):
        buf = C.create_string_buffer(bytes(msg), len(msg))
        check(lib().gc_view_request_apply(C.byref(self.r), buf, len(msg)))

    def view(self, width, height, precision=32):
        v = View()
        check(lib().gc_view_request_view(C.byref(self.r), width, height, precision, C.byref(v)))
        return v

    def tf(self):
        return np.ctypeslib.as_array(self.r.tf)[:4 * self.r.tf_entries].reshape(-1, 4).copy()


def set_ray_list_max(cache, n):
    """Resumed-ray launches of at most n rays march a group of lanes per ray (0: never)."""
    check(lib().gc_cache_set_ray_list_max(cache.h, n))


class LaunchLog:
    """Per-launch records of the following render_frame calls on a cache."""

    def __init__(self, cache, capacity=4096):
        self.cache = cache
        self.records = (LaunchRecord * capacity)()
        self.count = C.c_uint32(0)
        check(lib().gc_cache_set_launch_log(cache.h, self.records, capacity, C.byref(self.count)))

    def last_frame(self):
        """The last frame's launches as dicts."""
        out = []
        for r in self.records[:self.count.value]:
            out.append({"launch": r.launch, "kernel": ("lane", "group", "mop")[r.kernel], "list_rays": r.list_rays,
                        "group": r.group, "mapped_samples": r.stats.mapped_samples,
                        "brick_visits": r.stats.brick_visits, "rays": r.stats.pixels,
                        "unmapped_rays": r.stats.unmapped_rays, "kernel_ms": r.kernel_ms,
                        "bricks_loaded": r.bricks_loaded, "handed_over": r.handed_over})
        return out

    def close(self):
        if self.cache is not None and self.cache.h:
            check(lib().gc_cache_set_launch_log(self.cache.h, None, 0, None))
        self.cache = None


def set_ray_group(cache, lanes):
    """Lanes per ray of the group-per-ray march (4, 8, 16, 64; 0: from the list length)."""
    check(lib().gc_cache_set_ray_group(cache.h, lanes))


def set_ray_handoff(cache, samples, lanes=0, exited_pct=0):
    """Hand rays that have consumed `samples` samples in a lane-per-ray launch of
    render_frame over to a group march of `lanes` lanes per ray (0: 16) within
    the launch, once exited_pct percent of the launch's waves have run out of
    rays (0: from the start); samples = 0 turns it off.  The results are the
    same."""
    check(lib().gc_cache_set_ray_handoff(cache.h, samples, lanes, exited_pct))


def get_ray_handoff(cache):
    """(samples, lanes, exited_pct) of the cache's hand-over setting."""
    v = [C.c_uint32(0) for _ in range(3)]
    check(lib().gc_cache_get_ray_handoff(cache.h, *[C.byref(x) for x in v]))
    return tuple(x.value for x in v)


def set_max_requests(cache, n):
    check(lib().gc_cache_set_max_requests(cache.h, n))


def unpack_rgba(px):
    px = np.asarray(px, np.uint32)
    return np.stack([(px >> s) & 0xFF for s in (0, 8, 16, 24)], axis=-1).astype(np.int32)
