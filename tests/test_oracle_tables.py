"""Oracle pinning (CPU): the multi-level page table -- page-table caches
between the MRPD and the brick pool (CacheManager.hpp:151-236,
CacheManagerKernels.hpp:111-357, K_CacheDevice.hpp:216-334,
K_CacheManager.hpp:164-361).

Pins:
  * geometry: realNumberOfEntries / virtualizedNumberOfEntries of one- and
    two-table hierarchies against hand-computed values;
  * transparency: when the table block sizes divide the bricks per axis of
    every level (big bricks of that size), the padding is nil, the FP32
    positions are those of the MRPD-only cache, and a table cache that holds
    every page block changes nothing the data cache does.  With the same
    request stream the LRU, usage, pool, request buffer, frames and probes
    equal the MRPD-only cache's at every step;
  * a hand-traced eviction: a one-block table cache moving between two page
    blocks clears the old MRPD cell; a stale copy of a brick whose block was
    evicted is later evicted without touching the brick's newer mapping;
  * invariants under random request streams with small table caches
    (evictions at every level): every Mapped probe reads the level volume's
    2x2x2 box at its voxel, no two entries name one block or slot, table
    LRUs stay permutations of their blocks and record their blocks' owners.
"""
import ctypes as C

import numpy as np
import pytest

import gcvol as G
import oracle_lib as O
from scenes import Scene, orbit, tf_table


def _cfg(real, tables, bs=8, cov=1, cache=(4, 4, 1)):
    return O.make_cfg(real, real, bs, cov, 1, cache, tables=tables)


def test_geometry_hand_values(oracle):
    # level 0 96x64x40, level 1 48x32x20, bricks of 8 with covering 1
    real = [(96, 64, 40), (48, 32, 20)]
    g = O.geometry(_cfg(real, [((2, 2, 2), (4, 4, 2))]))
    assert g.nb_tables == 1
    # bricks: ceil(real / 6)
    assert [list(g.level_dim[l]) for l in range(2)] == [[16, 11, 7], [8, 6, 4]]
    # MRPD: ceil(bricks / block)
    assert [list(g.top_dim[l]) for l in range(2)] == [[4, 3, 4], [2, 2, 2]]
    assert list(g.top_coord)[:2] == [0, 4]
    assert list(g.top_cells_dim) == [6, 3, 4]
    assert list(g.mrpd_dim) == [24, 11, 7]
    t = g.table[0]
    assert list(t.block) == [4, 4, 2] and list(t.blocks) == [2, 2, 2] and t.nblocks == 8
    assert list(t.surf) == [8, 8, 4]
    # virtualized: MRPD * block, then * (bs - 2c)
    assert [list(t.virt[l]) for l in range(2)] == [[16, 12, 8], [8, 8, 4]]
    assert [list(t.real[l]) for l in range(2)] == [[16, 11, 7], [8, 6, 4]]
    assert [list(g.virt_vox[l]) for l in range(2)] == [[96, 72, 48], [48, 48, 24]]
    assert g.rdv[0][1] == np.float32(64) / np.float32(72)

    # two tables: CacheSize[1] (block 4,4,2) next to the pool, CacheSize[2]
    # (block 2,2,2) next to the MRPD
    g2 = O.geometry(_cfg(real[:1], [((2, 2, 2), (4, 4, 2)), ((3, 3, 3), (2, 2, 2))]))
    assert g2.nb_tables == 2
    assert list(g2.top_dim[0]) == [2, 2, 2]
    t0, t1 = g2.table[0], g2.table[1]
    assert list(t0.block) == [2, 2, 2] and list(t0.blocks) == [3, 3, 3]
    assert list(t0.virt[0]) == [4, 4, 4] and list(t0.real[0]) == [4, 3, 4]
    assert list(t1.block) == [4, 4, 2] and list(t1.blocks) == [2, 2, 2]
    assert list(t1.virt[0]) == [16, 16, 8] and list(t1.real[0]) == [16, 11, 7]
    assert list(g2.virt_vox[0]) == [96, 96, 48]

    # no table: the MRPD is the brick grid
    g0 = O.geometry(_cfg(real, []))
    assert g0.nb_tables == 0
    assert [list(g0.top_dim[l]) for l in range(2)] == [list(g0.level_dim[l]) for l in range(2)]
    assert list(g0.top_cells_dim) == list(g0.mrpd_dim)
    assert [list(g0.virt_vox[l]) for l in range(2)] == [[96, 66, 42], [48, 36, 24]]


def _full_tables(sc_kw, blocks):
    """A scene with tables of the given block sizes (config order) sized to
    hold every page block."""
    probe = Scene(tables=[((1, 1, 1), b) for b in blocks], **sc_kw)
    g = O.geometry(probe.ocfg)
    T = len(blocks)
    sizes = []
    for k in range(T):
        t = g.table[T - 1 - k]
        n = 0
        for l in range(probe.vc.nb_levels):
            n += int(np.prod([-(-int(t.real[l][a]) // int(t.block[a])) for a in range(3)]))
        sizes.append((n, 1, 1))
    return Scene(tables=[(sizes[k], blocks[k]) for k in range(T)], **sc_kw)


def _drive(caches, view, steps=60, probe_every=0, rng=None):
    """Render + update the caches in lockstep, yielding after every step."""
    W, H = view.width, view.height
    bufs = [(np.zeros(W * H, np.float32), np.zeros(W * H * 4), np.zeros(W * H, np.uint32)) for _ in caches]
    for step in range(steps):
        stats = [c.render(view, *b, precision=64) for c, b in zip(caches, bufs)]
        yield step, stats, bufs
        dones = [c.update() for c in caches]
        assert len(set(dones)) == 1
        if dones[0]:
            return


@pytest.mark.parametrize("blocks,bb", [([(2, 2, 2)], (2, 2, 2)), ([(2, 2, 1), (2, 1, 2)], (4, 4, 4)),
                                       ([(2, 1, 1), (1, 2, 1), (1, 1, 2)], (2, 2, 2))])
def test_transparent_tables_change_nothing(oracle, blocks, bb):
    kw = dict(dims=(72, 56, 40), bs=8, cov=1, bytes_=1, cache_size=(5, 4, 3), max_requests=17, seed=11,
              big_brick=bb, nb_blobs=5, sigma=(3.0, 9.0))
    flat = Scene(**kw)
    tab = _full_tables(kw, blocks)
    a, b = flat.oracle_cache(), tab.oracle_cache()
    # no padding: identical FP32 positions
    assert np.array_equal(np.ctypeslib.as_array(a.g.virt_vox), np.ctypeslib.as_array(b.g.virt_vox))
    view = O.make_view(flat.ocfg, 40, 32, orbit(3), tf_table(seed=2))
    rng = np.random.default_rng(0)
    loads = 0
    for step, stats, bufs in _drive([a, b], view, steps=80):
        assert np.array_equal(stats[0], stats[1]), step
        for x, y in zip(bufs[0], bufs[1]):
            assert np.array_equal(x, y), step
        assert np.array_equal(a.lru(), b.lru()), step
        assert np.array_equal(a.usage(), b.usage()), step
        assert np.array_equal(a.request(), b.request()), step
        assert np.array_equal(a.pool(), b.pool()), step
        loads += int(b.unreferenced() == 0)
        # probes: both API modes, same answers and side effects
        for _ in range(20):
            lvl = int(rng.integers(0, flat.vc.nb_levels))
            pos = (rng.random(3) * 0.99).astype(np.float32)
            mode = int(rng.integers(0, 2))
            assert a.probe(mode, lvl, pos) == b.probe(mode, lvl, pos)
        assert np.array_equal(a.request(), b.request())
        assert np.array_equal(a.usage(), b.usage())
    assert b.unreferenced() == 0
    assert (np.ctypeslib.as_array(a.mrpd())[:, 3] == 1).sum() > 0


def _req(c, level, b):
    g = c.g
    cell = g.level_coord[level] + b[0] + g.mrpd_dim[0] * (b[1] + g.mrpd_dim[1] * b[2])
    c.request()[cell] = c.ts


def _vol(c, level, b):
    out = np.zeros(3, np.float32)
    O.lib().or_volume_position(C.byref(c.g), level, O.ptr(np.asarray(b, np.uint32)), O.ptr(out))
    return out


def test_hand_traced_eviction(oracle):
    # one level of 4x2x2 bricks (24x12x12 voxels); page blocks of 2x2x2
    # bricks: two MRPD cells; a table cache of ONE block; 16 brick slots
    cfg = O.make_cfg([(24, 12, 12)], [(24, 12, 12)], 8, 1, 1, (4, 4, 1), tables=[((1, 1, 1), (2, 2, 2))])
    lv = [np.arange(24 * 12 * 12, dtype=np.uint32).reshape(12, 12, 24).astype(np.uint8)]
    c = O.Cache(cfg, lv)
    assert list(c.g.top_cells_dim) == [2, 1, 1]
    origin = lambda s: [(s % 4) * 8, (s // 4) * 8, 0]   # noqa: E731

    def load(bricks, stamp=None):
        for b in bricks:
            _req(c, 0, b)
        assert not c.update()              # request list handed over
        if stamp is not None:
            c.usage()[stamp] = c.ts        # used in the frame before the load
        assert not c.update()              # load

    # A: brick (0,0,0) -> slot 15 (LRU tail), MRPD cell 0 -> the block
    load([(0, 0, 0)])
    m, t = c.mrpd(), c.table(0)
    assert m[0].tolist() == [0, 0, 0, 1] and m[1].tolist() == [0, 0, 0, 0]
    assert t[0].tolist() == origin(15) + [1] and not t[1:].any()
    rec = c.table_lru(0)[0]
    assert rec[3] == 0 and np.array_equal(rec[4:7].view(np.float32), _vol(c, 0, (0, 0, 0)))
    # B: slot 15 used in the last frame; brick (2,0,0) -> slot 14; its page
    # block needs the table's only block: cell 0 is cleared, the block reset
    load([(2, 0, 0)], stamp=[15])
    m, t = c.mrpd(), c.table(0)
    assert m[0].tolist() == [0, 0, 0, 0] and m[1].tolist() == [0, 0, 0, 1]
    assert t[0].tolist() == origin(14) + [1] and not t[1:].any()
    assert np.array_equal(c.table_lru(0)[0, 4:7].view(np.float32), _vol(c, 0, (2, 0, 0)))
    # C: brick (0,0,0) again -> slot 14 (tail; brick (2,0,0) dereferenced),
    # the block moves back to cell 0.  Slot 15 still holds a stale copy.
    load([(0, 0, 0)])
    m, t = c.mrpd(), c.table(0)
    assert m[0].tolist() == [0, 0, 0, 1] and m[1].tolist() == [0, 0, 0, 0]
    assert t[0].tolist() == origin(14) + [1] and not t[1:].any()
    # D: every slot but 15 used: the stale copy in slot 15 is the victim of
    # brick (1,0,0), which lands in the mapped block without a table victim;
    # the stale copy's eviction leaves the live mapping of (0,0,0) alone
    load([(1, 0, 0)], stamp=[s for s in range(16) if s != 15])
    m, t = c.mrpd(), c.table(0)
    assert m[0].tolist() == [0, 0, 0, 1]
    assert t[0].tolist() == origin(14) + [1]
    assert t[1].tolist() == origin(15) + [1]
    assert not t[2:].any()
    # the two mapped bricks read their own voxels
    for b, s in (((0, 0, 0), 14), ((1, 0, 0), 15)):
        st, s8, bb, v = c.sample(0, _vol(c, 0, b) + np.float32(0.01))
        assert st == 0 and list(bb) == list(b)


def _box_truth(level, gv):
    Z, Y, X = level.shape
    tot = 0
    for dz in (-1, 0):
        for dy in (-1, 0):
            for dx in (-1, 0):
                x, y, z = gv[0] + dx, gv[1] + dy, gv[2] + dz
                if 0 <= x < X and 0 <= y < Y and 0 <= z < Z:
                    tot += int(level[z, y, x])
    return tot


def _check_tables(sc, c, rng):
    g = c.g
    T = c.nb_tables
    inner = sc.bs - 2 * sc.cov
    # table LRUs: permutations of their blocks
    for i in range(T):
        t = g.table[i]
        lru = c.table_lru(i)
        org = lru[:, :3]
        assert (org % np.array(t.block) == 0).all()
        ids = org[:, 0] // t.block[0] + t.blocks[0] * (org[:, 1] // t.block[1] + t.blocks[1] * (org[:, 2] // t.block[2]))
        assert sorted(ids.tolist()) == list(range(t.nblocks))
    # Unmapped entries are zero; no block or slot is named twice by the
    # entries a walk can reach.  (A table block whose parent block was
    # evicted is unreachable and keeps its entries until it is reused, which
    # resets it first, as in the reference.)
    for arr in [c.mrpd()] + [c.table(i) for i in range(T)]:
        assert (arr[arr[:, 3] == 0] == 0).all()
    named = c.mrpd()[c.mrpd()[:, 3] == 1][:, :3]
    for i in range(T + 1):
        assert len({tuple(r) for r in named}) == len(named), i
        if i == T:
            break
        t, tab = g.table[i], c.table(i)
        S = np.array(t.surf)
        nxt = []
        for o in named:
            idx = [(o[0] + x) + S[0] * ((o[1] + y) + S[1] * (o[2] + z))
                   for z in range(t.block[2]) for y in range(t.block[1]) for x in range(t.block[0])]
            e = tab[idx]
            nxt.append(e[e[:, 3] == 1][:, :3])
        named = np.concatenate(nxt) if nxt else np.zeros((0, 3), np.uint32)
    # MRPD entries name blocks whose owner position maps back to the cell
    if T:
        lru0 = c.table_lru(0)
        by_origin = {tuple(r[:3]): r for r in lru0}
        for cell in np.nonzero(c.mrpd()[:, 3] == 1)[0]:
            r = by_origin[tuple(c.mrpd()[cell, :3])]
            pos = np.ascontiguousarray(r[4:7].view(np.float32))
            assert int(O.lib().or_top_index(C.byref(g), int(r[3]), O.ptr(pos))) == cell
    # every Mapped sample reads its brick
    for l in range(sc.vc.nb_levels):
        lv = sc.levels[l]
        for _ in range(60):
            pos = (rng.random(3) * 0.99).astype(np.float32)
            st, s8, b, v = c.sample(l, pos)
            if st != 0:
                continue
            gv = (b.astype(np.int64) * inner + v.astype(np.int64) % inner).tolist()
            assert s8 == _box_truth(lv, gv)


@pytest.mark.parametrize("tables", [
    [((2, 1, 1), (2, 2, 2))],
    [((3, 2, 1), (2, 2, 1))],
    [((3, 3, 2), (2, 2, 2)), ((2, 1, 1), (2, 1, 2))],
    [((4, 3, 3), (2, 2, 1)), ((3, 2, 2), (1, 2, 2)), ((2, 2, 1), (2, 1, 1))],
])
def test_random_stream_invariants(oracle, tables):
    sc = Scene((64, 48, 40), 8, 1, 1, cache_size=(5, 3, 3), max_requests=23, seed=5, nb_blobs=4,
               sigma=(3.0, 8.0), tables=tables)
    c = sc.oracle_cache()
    g = c.g
    rng = np.random.default_rng(len(tables) * 7 + tables[0][0][0])
    mapped_seen = 0
    for step in range(40):
        ts = c.ts
        req = c.request()
        for _ in range(rng.integers(0, 30)):
            l = rng.integers(0, sc.vc.nb_levels)
            b = [rng.integers(0, g.level_dim[l][a]) for a in range(3)]
            req[(g.level_coord[l] + b[0]) + g.mrpd_dim[0] * (b[1] + g.mrpd_dim[1] * b[2])] = ts
        c.usage()[rng.random(c.nslots) < 0.2] = ts
        before = []
        for i in range(c.nb_tables):
            u = c.table_usage(i)
            u[rng.random(len(u)) < 0.3] = ts
            t = g.table[i]
            lru = c.table_lru(i).copy()
            blk = lru[:, 0] // t.block[0] + t.blocks[0] * (lru[:, 1] // t.block[1] + t.blocks[1] * (lru[:, 2] // t.block[2]))
            used = u[blk] == ts
            before.append(np.concatenate([lru[used], lru[~used]]))
        c.update()
        # each table's LRU: the blocks used in the frame first, both groups
        # in their old order (a load only rewrites the records' owners)
        for i in range(c.nb_tables):
            assert np.array_equal(c.table_lru(i)[:, :3], before[i][:, :3]), (step, i)
        _check_tables(sc, c, rng)
        mapped_seen += int((c.table(c.nb_tables - 1)[:, 3] == 1).sum())
    assert mapped_seen > 0


def test_rendered_to_completion_with_tables(oracle):
    # tables large enough for the view's working set: the progressive frame
    # equals the fully resident MRPD-only frame (both FP64, resume rounding)
    kw = dict(dims=(48, 48, 48), bs=8, cov=1, bytes_=1, seed=7, nb_blobs=5, sigma=(4.0, 10.0))
    flat = Scene(**kw)
    res = flat.oracle_cache()
    res.load_all()
    tf = tf_table(seed=3)
    W, H = 48, 40
    view = O.make_view(flat.ocfg, W, H, orbit(4), tf)
    px = np.zeros(W * H, np.uint32)
    res.render(view, np.zeros(W * H, np.float32), np.zeros(W * H * 4), px, precision=64)
    tab = Scene(tables=[((8, 4, 4), (2, 2, 2))], **kw)
    oc = tab.oracle_cache()
    tb, cs, px2 = np.zeros(W * H, np.float32), np.zeros(W * H * 4), np.zeros(W * H, np.uint32)
    for _ in range(200):
        st = oc.render(view, tb, cs, px2, precision=64)
        if oc.update():
            break
    assert st[3] == 0
    err = np.abs(G.unpack_rgba(px2) - G.unpack_rgba(px))
    assert err.mean() < 0.05
    assert (err > 2).mean() < 0.002


@pytest.mark.parametrize("tables", [[], [((2, 2, 2), (4, 4, 2))], [((3, 2, 1), (3, 2, 2))],
                                    [((2, 2, 2), (4, 4, 2)), ((3, 3, 3), (2, 2, 2))],
                                    [((2, 2, 2), (2, 2, 2)), ((2, 2, 2), (2, 1, 2)), ((1, 2, 2), (3, 2, 2))]])
def test_library_geometry_matches_oracle(oracle, tables):
    """gc_geometry_compute (host side of the library, no device) against the
    oracle's hierarchy geometry."""
    sc = Scene((150, 97, 61), 8, 1, 1, cache_size=(4, 4, 4), tables=tables, nb_blobs=0)
    lg = G.geometry_compute(sc.vc, sc.cc)
    og = O.geometry(sc.ocfg)
    L = sc.vc.nb_levels
    A = np.ctypeslib.as_array
    for f in ("level_dim", "virt_vox", "ior", "rdv", "top_dim"):
        assert np.array_equal(A(getattr(lg, f))[:L], A(getattr(og, f))[:L]), f
    for f in ("level_coord", "top_coord"):
        assert np.array_equal(A(getattr(lg, f))[:L], A(getattr(og, f))[:L]), f
    for f in ("mrpd_dim", "top_cells_dim", "one_over"):
        assert np.array_equal(A(getattr(lg, f)), A(getattr(og, f))), f
    assert lg.nb_tables == og.nb_tables == len(tables)
    assert lg.top_cells == int(np.prod(A(og.top_cells_dim).astype(np.int64)))
    for i in range(len(tables)):
        lt, ot = lg.table[i], og.table[i]
        for f in ("block", "blocks", "surf"):
            assert np.array_equal(A(getattr(lt, f)), A(getattr(ot, f))), (i, f)
        assert lt.nblocks == ot.nblocks
        for f in ("virt", "real"):
            assert np.array_equal(A(getattr(lt, f))[:L], A(getattr(ot, f))[:L]), (i, f)
