"""Oracle pinning (CPU): LRU update, miss list, request handling and loads.

Pins:
  * LRU update == a stable sort of the LRU by "not used this frame";
  * miss list == np.nonzero over [0, cells-1) (last cell excluded), first
    `cap` entries, decoded with unravel_index + a search over level offsets;
  * handler == dedup by big brick, non-empty first (stable);
  * after every update of a random trace: page directory <-> LRU bijection on
    mapped entries, every mapped slot holds its source brick, the LRU stays a
    permutation of the slots, Empty bricks map to (0,0,0,2), victims are the
    last non-empty-count LRU entries.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from scenes import Scene


def lru_slots(lru, bs, cs):
    cp = lru[:, :3] // bs
    return cp[:, 0] + cs[0] * (cp[:, 1] + cs[1] * cp[:, 2])


def test_lru_update_is_stable_partition(oracle):
    sc = Scene((32, 32, 32), 8, 1, 1, cache_size=(5, 4, 3), seed=1, nb_blobs=3, sigma=(3.0, 6.0))
    c = sc.oracle_cache()
    rng = np.random.default_rng(0)
    for step in range(6):
        before = c.lru().copy()
        slots = lru_slots(before, 8, sc.cache_size)
        use = rng.random(c.nslots) < 0.3
        u = c.usage()
        u[:] = np.where(use, c.ts, rng.integers(0, max(c.ts, 1), c.nslots))
        used = u[slots] == c.ts
        expect = before[sorted(range(c.nslots), key=lambda i: (not used[i], i))]
        c.lru_update()
        assert np.array_equal(c.lru(), expect)
        # advance the timestamp through a no-request update
        c.update()


def test_request_list_last_cell_and_cap(oracle):
    sc = Scene((64, 48, 40), 8, 1, 1, cache_size=(4, 4, 4), seed=2, nb_blobs=3, sigma=(3.0, 6.0))
    c = sc.oracle_cache()
    g = c.g
    rng = np.random.default_rng(1)
    req = c.request()
    ts = c.ts
    req[:] = np.where(rng.random(c.cells) < 0.05, ts, ts - 1)
    req[-1] = ts                                   # never collected
    for cap in (1, 7, 50, 100000):
        got = c.request_list(cap)
        idx = np.nonzero(req[: c.cells - 1] == ts)[0][:cap]
        sx = int(g.mrpd_dim[0])
        sy = int(g.level_dim[0][1])
        z, rem = np.divmod(idx, sx * sy)
        y, x = np.divmod(rem, sx)
        ends = np.cumsum([g.level_dim[l][0] for l in range(sc.vc.nb_levels)])
        level = np.searchsorted(ends, x, side="right")
        starts = np.concatenate([[0], ends[:-1]])
        exp = np.stack([x - starts[level], y, z, level], 1).astype(np.uint32)
        assert np.array_equal(got, exp)
    assert len(c.request_list(100000)) == int((req[:-1] == ts).sum())


def _check_invariants(sc, c, levels):
    g = c.g
    bs, cs = sc.bs, sc.cache_size
    mrpd, lru, pool = c.mrpd(), c.lru(), c.pool()
    assert set(np.unique(mrpd[:, 3])) <= {0, 1, 2}
    # LRU is a permutation of the slots
    slots = lru_slots(lru, bs, cs)
    assert sorted(slots.tolist()) == list(range(c.nslots))
    assert (lru[:, :3] % bs == 0).all()
    vol = lru[:, 4:7].view(np.float32)
    import ctypes as C
    mapped_cells = np.nonzero(mrpd[:, 3] == 1)[0]
    by_origin = {tuple(r[:3]): i for i, r in enumerate(lru)}
    for cell in mapped_cells:
        e = mrpd[cell]
        i = by_origin[tuple(e[:3])]
        pos = np.ascontiguousarray(vol[i])
        back = sc_oracle_index(sc, c, int(lru[i, 3]), pos)
        assert back == cell, "LRU entry %d does not map back to its cell" % i
    # distinct mapped cells use distinct slots
    assert len({tuple(mrpd[cl, :3]) for cl in mapped_cells}) == len(mapped_cells)
    # mapped slots hold their source brick
    for cell in mapped_cells[:: max(1, len(mapped_cells) // 40)]:
        e = mrpd[cell]
        i = by_origin[tuple(e[:3])]
        level = int(lru[i, 3])
        x = cell % g.mrpd_dim[0]
        y = (cell // g.mrpd_dim[0]) % g.mrpd_dim[1]
        z = cell // (g.mrpd_dim[0] * g.mrpd_dim[1])
        b = (int(x - g.level_coord[level]), int(y), int(z))
        exp = sc_brick(sc, levels, level, b)
        got = pool[e[2]:e[2] + bs, e[1]:e[1] + bs, e[0]:e[0] + bs]
        assert np.array_equal(got, exp)
    empty = mrpd[mrpd[:, 3] == 2]
    assert (empty[:, :3] == 0).all()


def sc_oracle_index(sc, c, level, pos):
    import ctypes as C
    import oracle_lib as O
    return int(O.lib().or_mrpd_index(C.byref(c.g), level, O.ptr(pos)))


def sc_brick(sc, levels, level, b):
    import oracle_lib as O
    return O.brick_extract(levels[level], sc.bs, sc.cov, b)


@pytest.mark.parametrize("seed", [0, 1])
def test_random_trace_invariants(oracle, seed):
    sc = Scene((48, 40, 32), 8, 1, 1, cache_size=(4, 3, 3), seed=seed, max_requests=13, nb_blobs=4,
               sigma=(3.0, 8.0))
    c = sc.oracle_cache()
    rng = np.random.default_rng(seed)
    g = c.g
    bs = sc.bs
    cx, cy = sc.cache_size[0], sc.cache_size[1]
    loaded = 0
    will_load = False     # the last update handed a request to the handler
    for step in range(40):
        ts = c.ts
        # a frame: random requests on valid cells, usage on some mapped slots
        req = c.request()
        for _ in range(rng.integers(0, 25)):
            l = rng.integers(0, sc.vc.nb_levels)
            b = [rng.integers(0, g.level_dim[l][a]) for a in range(3)]
            cell = (g.level_coord[l] + b[0]) + g.mrpd_dim[0] * (b[1] + g.mrpd_dim[1] * b[2])
            if c.mrpd()[cell, 3] == 0:
                req[cell] = ts
        u = c.usage()
        u[rng.random(c.nslots) < 0.2] = ts
        pending = c.request_list(100000)
        lru_before = c.lru().copy()
        asked_before, nne_before = c.asked()
        usage = c.usage().copy()
        done = c.update()
        assert c.ts == ts + 1
        if done:
            assert len(pending) == 0
        _check_invariants(sc, c, sc.levels)
        # LRU update: the slots used in this frame first, each group in its
        # old order (CacheDevice.hpp:202-219)
        slot = lru_before[:, 0] // bs + cx * (lru_before[:, 1] // bs + cy * (lru_before[:, 2] // bs))
        used = usage[slot] == ts
        part = np.concatenate([lru_before[used], lru_before[~used]])
        after = c.lru()
        assert np.array_equal(after[:, :3], part[:, :3]), step
        if will_load and nne_before:
            # load: the victims are the last nne entries, in order, and take
            # the asked bricks' level and volume position
            # (CacheManager.hpp:511-512,554-569)
            N = c.nslots
            assert np.array_equal(after[:N - nne_before], part[:N - nne_before]), step
            for j in range(nne_before):
                a = asked_before[j]
                vp = np.zeros(3, np.float32)
                O.lib().or_volume_position(C.byref(c.g), int(a[3]), O.ptr(np.ascontiguousarray(a[:3], np.uint32)),
                                           O.ptr(vp))
                assert after[N - nne_before + j, 3] == a[3], (step, j)
                assert np.array_equal(after[N - nne_before + j, 4:7], vp.view(np.uint32)), (step, j)
            loaded += nne_before
            will_load = False
        else:
            assert np.array_equal(after, part), step
            will_load = not done
    assert loaded > 0


def test_handler_big_bricks_and_empty_first(oracle):
    # zero volume except one corner: with threshold 1 most big bricks are Empty
    dims = (32, 32, 32)
    vol = np.zeros((32, 32, 32), np.uint8)
    vol[:6, :6, :6] = 200
    import oracle_lib as O
    import gcvol as G
    vc = G.volume_cfg(dims, 8, 1, 1, (2, 2, 1))
    levels = [vol]
    for _ in range(1, vc.nb_levels):
        levels.append(O.mip_reduce(levels[-1]))
    cfg = O.make_cfg(vc.initial(), vc.real(), 8, 1, 1, (4, 4, 4), bb=(2, 2, 1), max_requests=50,
                     empty_threshold=1.0)
    c = O.Cache(cfg, levels)
    g = c.g
    req = c.request()
    ts = c.ts
    asked = [(0, 0, 0, 0), (1, 0, 0, 0), (3, 3, 3, 0), (0, 1, 0, 0), (2, 2, 2, 0), (1, 1, 1, 0)]
    for (x, y, z, l) in asked:
        req[(g.level_coord[l] + x) + g.mrpd_dim[0] * (y + g.mrpd_dim[1] * z)] = ts
    assert not c.update()
    got, nne = c.asked()
    # a big brick is Empty iff every voxel of it (covering included) is below
    # the threshold; non-empty small bricks come first, each group in request
    # (linear index) order
    lin = sorted(asked, key=lambda a: (a[2], a[1], a[0]))

    def big_empty(a):
        bb = O.brick_extract(vol, 8, 1, (a[0] // 2, a[1] // 2, a[2]), (2, 2, 1))
        return bool((bb < 1).all())

    ne = [a for a in lin if not big_empty(a)]
    em = [a for a in lin if big_empty(a)]
    assert ne and em
    assert nne == len(ne)
    assert [tuple(r) for r in got] == ne + em
    c.update()
    mrpd = c.mrpd()
    for (x, y, z, l) in em:
        cell = (g.level_coord[l] + x) + g.mrpd_dim[0] * (y + g.mrpd_dim[1] * z)
        assert mrpd[cell].tolist() == [0, 0, 0, 2]
    for (x, y, z, l) in ne:
        cell = (g.level_coord[l] + x) + g.mrpd_dim[0] * (y + g.mrpd_dim[1] * z)
        assert mrpd[cell, 3] == 1


def test_update_protocol(oracle):
    """update(): requests raised in frame ts are compacted at ts, loaded on the
    next update, and an update with no request reports done."""
    sc = Scene((32, 32, 32), 8, 1, 1, cache_size=(4, 4, 4), seed=3, nb_blobs=3, sigma=(3.0, 6.0))
    c = sc.oracle_cache()
    assert c.update()                       # nothing requested
    g = c.g
    ts = c.ts
    cell = g.level_coord[1] + 1 + g.mrpd_dim[0] * (2 + g.mrpd_dim[1] * 0)
    c.request()[cell] = ts
    assert not c.update()                   # request handled, not yet loaded
    assert c.mrpd()[cell, 3] == 0
    assert not c.update()                   # load
    assert c.mrpd()[cell, 3] == 1
    lru = c.lru()
    # the victim was the last LRU entry
    assert lru[-1, 3] == 1
    assert c.update()
    assert abs(c.completude() - 1.0 / c.nslots) < 1e-7
