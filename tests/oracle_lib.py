"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py only.  The oracle restates the reference rules on
the CPU; see oracle/gcvol_oracle.h for the rule citations.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")
MAXL = 16
MAXT = 3

u32 = C.c_uint32
f32 = C.c_float


class OrCfg(C.Structure):
    _fields_ = [
        ("brick_size", u32), ("covering", u32), ("bytes", u32),
        ("big_brick", u32 * 3), ("nb_levels", u32),
        ("initial", (u32 * 3) * MAXL), ("real", (u32 * 3) * MAXL),
        ("cache_size", u32 * 3), ("max_requests", u32), ("empty_threshold", f32),
        ("nb_tables", u32), ("table_cache", (u32 * 3) * MAXT), ("table_block", (u32 * 3) * MAXT),
    ]


class OrTableGeom(C.Structure):
    _fields_ = [
        ("block", u32 * 3), ("blocks", u32 * 3), ("surf", u32 * 3), ("nblocks", u32),
        ("virt", (u32 * 3) * MAXL), ("real", (u32 * 3) * MAXL),
    ]


class OrGeom(C.Structure):
    _fields_ = [
        ("level_dim", (u32 * 3) * MAXL), ("level_coord", u32 * MAXL),
        ("virt_vox", (u32 * 3) * MAXL), ("ior", (f32 * 3) * MAXL), ("rdv", (f32 * 3) * MAXL),
        ("one_over", f32 * 3), ("mrpd_dim", u32 * 3), ("pool_dim", u32 * 3), ("nslots", u32),
        ("nb_tables", u32), ("top_dim", (u32 * 3) * MAXL), ("top_coord", u32 * MAXL),
        ("top_cells_dim", u32 * 3), ("table", OrTableGeom * MAXT),
    ]


class OrLru(C.Structure):
    _fields_ = [("cache_pos", u32 * 3), ("level", u32), ("vol_pos", f32 * 3)]


class OrView(C.Structure):
    _fields_ = [
        ("width", u32), ("height", u32), ("inv_mv", f32 * 12),
        ("bbox_min", f32 * 3), ("bbox_max", f32 * 3), ("tstep", f32), ("alpha_stop", f32),
        ("nb_lods", u32), ("inv_levels_size", (f32 * 3) * MAXL), ("lod_step", f32 * MAXL),
        ("tf", C.POINTER(f32)), ("tf_size", u32),
    ]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.or_level_count.restype = u32
        L.or_level_count.argtypes = [P, u32]
        L.or_level_sizes.argtypes = [P, P, u32, P]
        L.or_init_meta.argtypes = [u32, P, u32, u32, P, P, P, P]
        L.or_pool_shape.argtypes = [C.c_uint64, C.c_uint64, P]
        L.or_geometry.argtypes = [C.POINTER(OrCfg), C.POINTER(OrGeom)]
        L.or_mip_reduce.argtypes = [P, P, u32, P, P]
        L.or_brick_extract.argtypes = [P, P, u32, u32, u32, P, P, P]
        L.or_brick_level.argtypes = [P, P, u32, u32, u32, u32, u32, P]
        L.or_cache_create.restype = P
        L.or_cache_create.argtypes = [C.POINTER(OrCfg), P]
        L.or_cache_destroy.argtypes = [P]
        L.or_cache_geom.restype = C.POINTER(OrGeom)
        L.or_cache_geom.argtypes = [P]
        for name in ("or_cache_mrpd", "or_cache_request", "or_cache_usage"):
            getattr(L, name).restype = C.POINTER(u32)
            getattr(L, name).argtypes = [P]
        L.or_cache_lru.restype = C.POINTER(OrLru)
        L.or_cache_lru.argtypes = [P]
        L.or_cache_pool.restype = P
        L.or_cache_pool.argtypes = [P]
        L.or_cache_timestamp.restype = u32
        L.or_cache_timestamp.argtypes = [P]
        L.or_cache_update.restype = C.c_int
        L.or_cache_update.argtypes = [P]
        L.or_lru_update.argtypes = [P]
        L.or_request_list.restype = u32
        L.or_request_list.argtypes = [P, P, u32]
        L.or_completude.restype = f32
        L.or_completude.argtypes = [P]
        L.or_asked.restype = u32
        L.or_asked.argtypes = [P, P, P]
        L.or_cache_load_all.argtypes = [P]
        L.or_volume_position.argtypes = [C.POINTER(OrGeom), u32, P, P]
        L.or_mrpd_index.restype = C.c_uint64
        L.or_mrpd_index.argtypes = [C.POINTER(OrGeom), u32, P]
        L.or_view_tables.argtypes = [C.POINTER(OrView), C.POINTER(OrCfg)]
        L.or_probe.restype = C.c_int
        L.or_probe.argtypes = [P, C.c_int, u32, P, C.POINTER(C.c_double)]
        L.or_compute_lod.restype = u32
        L.or_compute_lod.argtypes = [C.POINTER(OrView), P, P, f32, f32]
        L.or_tf_samples.argtypes = [P, u32, u32, P]
        L.or_render_dvr.argtypes = [P, C.POINTER(OrView), P, P, P, C.c_int, C.c_int, P]
        L.or_render_dvr_rows.argtypes = [P, C.POINTER(OrView), P, P, P, C.c_int, C.c_int, u32, u32, P]
        L.or_render_dvr_resume.argtypes = [P, C.POINTER(OrView), P, P, P, P, C.c_int, C.c_int, u32, u32, P]
        L.or_render_mop.argtypes = [P, C.POINTER(OrView), P, P, P, P, C.c_int, C.c_int, P]
        L.or_sample.restype = C.c_int
        L.or_sample.argtypes = [P, u32, P, C.POINTER(C.c_double), P, P]
        for name in ("or_cache_table", "or_cache_table_usage"):
            getattr(L, name).restype = C.POINTER(u32)
            getattr(L, name).argtypes = [P, u32]
        L.or_cache_table_lru.restype = C.POINTER(OrLru)
        L.or_cache_table_lru.argtypes = [P, u32]
        L.or_cache_unreferenced.restype = u32
        L.or_cache_unreferenced.argtypes = [P]
        L.or_top_index.restype = C.c_uint64
        L.or_top_index.argtypes = [C.POINTER(OrGeom), u32, P]
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


DT = {1: np.uint8, 2: np.uint16, 4: np.float32}


def level_count(dims, bs):
    return int(lib().or_level_count(ptr(np.asarray(dims, np.uint32)), bs))


def level_sizes(dims, nlev, ratio=(2, 2, 2)):
    out = np.zeros((nlev, 3), np.uint32)
    lib().or_level_sizes(ptr(np.asarray(dims, np.uint32)), ptr(np.asarray(ratio, np.uint32)), nlev, ptr(out))
    return out


def init_meta(levels, bs, cov, bb=(1, 1, 1)):
    levels = np.ascontiguousarray(levels, np.uint32)
    n = len(levels)
    nb = np.zeros((n, 3), np.uint32)
    nbb = np.zeros((n, 3), np.uint32)
    real = np.zeros((n, 3), np.uint32)
    lib().or_init_meta(n, ptr(levels), bs, cov, ptr(np.asarray(bb, np.uint32)), ptr(nb), ptr(nbb), ptr(real))
    return nb, nbb, real


def pool_shape(budget, brick_bytes):
    out = np.zeros(3, np.uint32)
    lib().or_pool_shape(budget, brick_bytes, ptr(out))
    return tuple(int(v) for v in out)


def mip_reduce(vol, ratio=(2, 2, 2)):
    """vol: numpy array indexed [z, y, x]."""
    vol = np.ascontiguousarray(vol)
    bytes_ = vol.dtype.itemsize
    dims = np.array([vol.shape[2], vol.shape[1], vol.shape[0]], np.uint32)
    r = np.asarray(ratio, np.uint32)
    out = np.zeros((dims[2] // r[2], dims[1] // r[1], dims[0] // r[0]), vol.dtype)
    lib().or_mip_reduce(ptr(vol), ptr(dims), bytes_, ptr(r), ptr(out))
    return out


def brick_extract(level, bs, cov, bbc, bb=(1, 1, 1)):
    level = np.ascontiguousarray(level)
    dims = np.array([level.shape[2], level.shape[1], level.shape[0]], np.uint32)
    bb = np.asarray(bb, np.uint32)
    out = np.zeros((bs * bb[2], bs * bb[1], bs * bb[0]), level.dtype)
    lib().or_brick_extract(ptr(level), ptr(dims), level.dtype.itemsize, bs, cov, ptr(bb),
                           ptr(np.asarray(bbc, np.uint32)), ptr(out))
    return out


def brick_level(level, bs, cov, b0, b1):
    """Bricks [b0, b1) of a level ([z,y,x] array), x-fastest brick order."""
    level = np.ascontiguousarray(level)
    dims = np.array([level.shape[2], level.shape[1], level.shape[0]], np.uint32)
    out = np.zeros((b1 - b0, bs, bs, bs), level.dtype)
    lib().or_brick_level(ptr(level), ptr(dims), level.dtype.itemsize, bs, cov, b0, b1, ptr(out))
    return out


def make_cfg(initial, real, bs, cov, bytes_, cache_size, bb=(1, 1, 1), max_requests=50,
             empty_threshold=0.0, tables=()):
    """tables: the page-table caches as (CacheSize[k+1], BlockSize[k+1])
    pairs in CacheConfiguration order (k = 0 next to the data cache)."""
    cfg = OrCfg()
    cfg.brick_size = bs
    cfg.covering = cov
    cfg.bytes = bytes_
    cfg.big_brick[:] = list(bb)
    cfg.nb_levels = len(initial)
    for l in range(len(initial)):
        cfg.initial[l][:] = [int(v) for v in initial[l]]
        cfg.real[l][:] = [int(v) for v in real[l]]
    cfg.cache_size[:] = list(cache_size)
    cfg.max_requests = max_requests
    cfg.empty_threshold = empty_threshold
    cfg.nb_tables = len(tables)
    for k, (cs, bl) in enumerate(tables):
        cfg.table_cache[k][:] = [int(v) for v in cs]
        cfg.table_block[k][:] = [int(v) for v in bl]
    return cfg


def geometry(cfg):
    g = OrGeom()
    lib().or_geometry(C.byref(cfg), C.byref(g))
    return g


class Cache:
    """The oracle CacheManager over a host pyramid (list of [z,y,x] arrays)."""

    def __init__(self, cfg, levels):
        self.cfg = cfg
        self.levels = [np.ascontiguousarray(l) for l in levels]
        arr = (C.c_void_p * len(self.levels))(*[l.ctypes.data for l in self.levels])
        self._arr = arr
        self.h = lib().or_cache_create(C.byref(cfg), arr)
        self.g = lib().or_cache_geom(self.h).contents
        self.cells = int(self.g.mrpd_dim[0]) * int(self.g.mrpd_dim[1]) * int(self.g.mrpd_dim[2])
        d = self.g.top_cells_dim
        self.top_cells = int(d[0]) * int(d[1]) * int(d[2])
        self.nslots = int(self.g.nslots)
        self.nb_tables = int(self.g.nb_tables)

    def close(self):
        if self.h:
            lib().or_cache_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def mrpd(self):
        p = lib().or_cache_mrpd(self.h)
        return np.ctypeslib.as_array(p, (self.top_cells, 4))

    def table(self, i):
        """Entries of table cache i (i = 0 next to the MRPD): (entries, 4),
        the surface x-fastest."""
        t = self.g.table[i]
        n = int(t.surf[0]) * int(t.surf[1]) * int(t.surf[2])
        return np.ctypeslib.as_array(lib().or_cache_table(self.h, i), (n, 4))

    def table_usage(self, i):
        return np.ctypeslib.as_array(lib().or_cache_table_usage(self.h, i), (int(self.g.table[i].nblocks),))

    def table_lru(self, i):
        p = C.cast(lib().or_cache_table_lru(self.h, i), C.POINTER(u32))
        return np.ctypeslib.as_array(p, (int(self.g.table[i].nblocks), 7))

    def unreferenced(self):
        return int(lib().or_cache_unreferenced(self.h))

    def request(self):
        return np.ctypeslib.as_array(lib().or_cache_request(self.h), (self.cells,))

    def usage(self):
        return np.ctypeslib.as_array(lib().or_cache_usage(self.h), (self.nslots,))

    def lru(self):
        """LRU as a (nslots, 7) uint32 array: cache_pos xyz, level, vol_pos bits."""
        p = C.cast(lib().or_cache_lru(self.h), C.POINTER(u32))
        return np.ctypeslib.as_array(p, (self.nslots, 7))

    def pool(self):
        d = self.g.pool_dim
        n = int(d[0]) * int(d[1]) * int(d[2])
        dt = DT[self.cfg.bytes]
        p = C.cast(lib().or_cache_pool(self.h), C.POINTER(np.ctypeslib.as_ctypes_type(dt)))
        return np.ctypeslib.as_array(p, (n,)).reshape(int(d[2]), int(d[1]), int(d[0]))

    @property
    def ts(self):
        return int(lib().or_cache_timestamp(self.h))

    def update(self):
        return bool(lib().or_cache_update(self.h))

    def lru_update(self):
        lib().or_lru_update(self.h)

    def request_list(self, cap):
        out = np.zeros((max(cap, 1), 4), np.uint32)
        n = lib().or_request_list(self.h, ptr(out), cap)
        return out[:n].copy()

    def asked(self):
        n = lib().or_asked(self.h, None, None)
        out = np.zeros((max(n, 1), 4), np.uint32)
        nne = u32(0)
        lib().or_asked(self.h, ptr(out), C.byref(nne))
        return out[:n].copy(), int(nne.value)

    def completude(self):
        return float(lib().or_completude(self.h))

    def load_all(self):
        lib().or_cache_load_all(self.h)

    def sample(self, level, pos):
        s = C.c_double(0)
        b = np.zeros(3, np.uint32)
        v = np.zeros(3, np.uint32)
        st = lib().or_sample(self.h, level, ptr(np.asarray(pos, np.float32)), C.byref(s), ptr(b), ptr(v))
        return st, s.value, b, v

    def probe(self, mode, level, pos):
        """Device API probe with side effects: (status, raw value)."""
        raw = C.c_double(0.0)
        st = lib().or_probe(self.h, mode, level, ptr(np.ascontiguousarray(pos, np.float32)), C.byref(raw))
        return int(st), raw.value

    def render(self, view, tbuf, cstate, pixels, precision=64, nthreads=1, rpos=None):
        """One launch; rpos (W*H*4 float32) selects exact continuation."""
        st = np.zeros(4, np.uint64)
        if rpos is None:
            lib().or_render_dvr(self.h, C.byref(view), ptr(tbuf), ptr(cstate), ptr(pixels), precision,
                                nthreads, ptr(st))
        else:
            assert rpos.dtype == np.float32 and rpos.size == view.width * view.height * 4
            lib().or_render_dvr_resume(self.h, C.byref(view), ptr(tbuf), ptr(cstate), ptr(rpos), ptr(pixels),
                                       precision, nthreads, 0, view.height, ptr(st))
        return st

    def render_mop(self, view, tbuf, cstate, pixels, precision=64, nthreads=1, rpos=None):
        """One maximum-projection launch (RayCastMOPImpl); cstate[4i] is the
        normalised maximum."""
        st = np.zeros(4, np.uint64)
        if rpos is not None:
            assert rpos.dtype == np.float32 and rpos.size == view.width * view.height * 4
        lib().or_render_mop(self.h, C.byref(view), ptr(tbuf), ptr(cstate), None if rpos is None else ptr(rpos),
                            ptr(pixels), precision, nthreads, ptr(st))
        return st


def tf_samples(knots, n):
    knots = np.ascontiguousarray(knots, np.float32)
    out = np.zeros((n, 4), np.float32)
    lib().or_tf_samples(ptr(knots), len(knots), n, ptr(out))
    return out


def make_view(cfg, width, height, inv_mv, tf, alpha_stop=0.95, bbox=((-1, -1, -1), (1, 1, 1))):
    v = OrView()
    lib().or_view_tables(C.byref(v), C.byref(cfg))
    v.width = width
    v.height = height
    v.inv_mv[:] = [float(x) for x in np.asarray(inv_mv, np.float32).reshape(12)]
    v.bbox_min[:] = list(bbox[0])
    v.bbox_max[:] = list(bbox[1])
    v.alpha_stop = alpha_stop
    tf = np.ascontiguousarray(tf, np.float32)
    v._tf_keep = tf
    v.tf = tf.ctypes.data_as(C.POINTER(f32))
    v.tf_size = len(tf)
    return v
