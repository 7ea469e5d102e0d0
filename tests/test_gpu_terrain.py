"""The terrain volume executor (mcgpu_terrain_*, csrc/mcg_terrain.*) on the GPU against libmcgen's own CPU
density path (oracle/_ref/libmcgen_ref.so) and the restatement (oracle/mcgterrain.c), bit for bit:
final_density per chunk, the ore-vein patch, and the cache cells the CPU would hold after the density
(which candidate volume, and its values).  Programs come from libmcgen's own compiler and exporter over
the worlds of tests/terrain_worlds.py."""
import ctypes as C

import numpy as np
import pytest

import mcgpu
import terrain as T
import terrain_worlds as TW

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not T.RefTerrain.available(), reason="libmcgen reference build absent (make -C oracle ref)")]

SEED = -4172144997902289642
CHUNKS = [(0, 0), (-3, 5), (70001, -65535), (1, -1), (-1, 1), (12, -7), (-200000, 3), (5, 5)]


def same_bits(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return a.shape == b.shape and bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))))


def kinds_setup(seed=SEED):
    noises, funcs, fd, pre = TW.kinds_world()
    dv, rv, gv = TW.vein_roots()
    R = T.RefTerrain(seed, noises, funcs)
    r, p, d, ri, gp = (R.compile(x) for x in (fd, pre, dv, rv, gv))
    cids = sorted(set(R.cache_ids(p) + R.cache_ids(r)))
    W = T.terrain_world(R, r, veins=[(d, ri, gp, 11, 12, 13, 0.02)], pre=p, cache_ids=cids)
    return R, W, r, p


def check_cells(info, which, cells, ref_cells, cx, cz):
    for k, (cid, cands, cap, off) in enumerate(info):
        st, ch, v9, vals = ref_cells[k]
        j = int(which[k])
        if not (st and ch):
            assert j == -1, (k, j)
            continue
        assert j >= 0, (k, j)
        vol9, xr, zr, n = cands[j]
        absolute = (vol9[0], vol9[1], vol9[2], vol9[3] + (cx * 16 if xr else 0), vol9[4], vol9[5] + (cz * 16 if zr else 0),
                    vol9[6], vol9[7], vol9[8])
        assert absolute == v9, (k, absolute, v9)
        assert same_bits(cells[off:off + n], vals)


def test_kinds_world_matches_libmcgen_and_restatement():
    print("terrain: kinds world, %d chunks" % len(CHUNKS), flush=True)
    R, W, r, p = kinds_setup()
    G = mcgpu.TerrainWorld(W, max_chunks=5)
    assert G.max_chunks == 5
    info = G.cells()
    assert [c[0] for c in info] == [W.cache_ids[k] for k in range(W.ncache_ids)]
    for lo in range(0, len(CHUNKS), G.max_chunks):
        part = CHUNKS[lo:lo + G.max_chunks]
        dens, veins, cells, which = G.batch([c[0] for c in part], [c[1] for c in part], cells=True)
        for i, (cx, cz) in enumerate(part):
            rd, rcells = T.ref_chunk(R, W, r, cx, cz, pre=p)
            od, ov, ocells = T.chunk(W, cx, cz)
            assert same_bits(dens[i], rd), (cx, cz)
            assert same_bits(od, rd)
            assert np.array_equal(veins[i], ov), (cx, cz)
            check_cells(info, which[i], cells[i], rcells, cx, cz)


def test_beardifier_volumes():
    """The Beardifier's values come from the caller per chunk; chunks without one read the default 0."""
    R, W, r, p = kinds_setup(7)

    @mcgpu.McgBeardFn
    def beard(ud, chunk, vol, out):
        if chunk % 2:
            return 0
        sx, sy, sz, x0, y0, z0, dx, dy, dz = (vol[i] for i in range(9))
        iz, ix, iy = np.meshgrid(np.arange(sz), np.arange(sx), np.arange(sy), indexing="ij")
        v = (0.001 * (x0 + ix * dx) - 0.002 * (y0 + iy * dy) + 0.0005 * (z0 + iz * dz)).astype(np.float32).ravel()
        C.memmove(out, v.ctypes.data, 4 * v.size)
        return 1

    G = mcgpu.TerrainWorld(W, max_chunks=4)
    part = CHUNKS[:4]
    dens, _, _, _ = G.batch([c[0] for c in part], [c[1] for c in part], beard=beard)
    for i, (cx, cz) in enumerate(part):
        od, _, _ = T.chunk(W, cx, cz, index=i, beard=beard)
        assert same_bits(dens[i], od), (i, cx, cz)
    rd, _ = T.ref_chunk(R, W, r, part[1][0], part[1][1], pre=p)         # chunk 1: no Beardifier = libmcgen's default
    assert same_bits(dens[1], rd)


def test_overworld_like_density():
    print("terrain: overworld-like world", flush=True)
    noises, funcs, fd = TW.overworld_like()
    for seed in (0, SEED):
        R = T.RefTerrain(seed, noises, funcs)
        r = R.compile(fd)
        W = T.terrain_world(R, r)
        G = mcgpu.TerrainWorld(W, max_chunks=len(CHUNKS))
        dens, veins, _, _ = G.batch([c[0] for c in CHUNKS], [c[1] for c in CHUNKS])
        assert veins is None
        for i, (cx, cz) in enumerate(CHUNKS):
            rd, _ = T.ref_chunk(R, W, r, cx, cz)
            assert same_bits(dens[i], rd), (seed, cx, cz)


def test_refusals():
    R = T.RefTerrain(1, {}, {})
    r = R.compile({"type": "log", "input": {"type": "add", "left": 2.0, "right": {"type": "gradient", "axis": "y",
                   "from_coordinate": 0, "to_coordinate": 10, "from_value": 0.0, "to_value": 1.0}}})
    W = T.terrain_world(R, r)
    with pytest.raises(mcgpu.PluginError, match="not covered"):
        mcgpu.TerrainWorld(W)
    r2 = R.compile({"type": "gradient", "axis": "x", "from_coordinate": 0, "to_coordinate": 16, "from_value": 0.0,
                    "to_value": 1.0})
    W2 = T.terrain_world(R, r2)
    G = mcgpu.TerrainWorld(W2, max_chunks=2)
    with pytest.raises(mcgpu.PluginError, match="bad batch"):
        G.batch([0, 1, 2], [0, 0, 0])
    dens, _, _, _ = G.batch([3], [-4])
    x = 48 + np.arange(16, dtype=np.float32)
    assert np.array_equal(dens[0].reshape(16, 16, W2.nh)[0, :, 0], np.clip(x, 0, 16) / 16)


def test_engine_terrain_world_matches_restatement():
    """The engine's stand-in Overworld terrain world (mcb_plugin_terrain_world): GPU batches against the
    restatement on the same program; a device-resident batch in between leaves the next host batch equal."""
    import mcbiome
    eng = mcbiome.Engine("26.3", "overworld")
    chunks = CHUNKS[:6]
    cx, cz = [c[0] for c in chunks], [c[1] for c in chunks]
    for seed in (0, -1, SEED):
        print("terrain: engine world, seed %d" % seed, flush=True)
        tw = eng.terrain_world(seed)
        G = mcgpu.TerrainWorld(tw.world, max_chunks=16)
        dens, _, _, _ = G.batch(cx, cz)
        for i, (x, z) in enumerate(chunks):
            od, _, _ = T.chunk(tw.world, x, z)
            assert same_bits(dens[i], od), (seed, x, z)
        assert G.batch_device(cx, cz)
        assert same_bits(G.batch(cx, cz)[0], dens)
        G.close()
        tw.close()


def test_worlds_reuse_device_kits():
    """Terrain worlds made one after another take over the stream and buffers of freed ones on the device
    (grown when a world needs more, kept when it needs less): the kinds world (veins, cache cells) and the
    engine's stand-in world in turn with batch sizes 2, 16, 1 and 8, two of them alive at once, each batch
    equal to the restatement."""
    import mcbiome
    eng = mcbiome.Engine("26.3", "overworld")
    R, W, r, p = kinds_setup()
    chunks = CHUNKS[:4]
    cx, cz = [c[0] for c in chunks], [c[1] for c in chunks]
    want_kinds = [T.chunk(W, x, z) for x, z in chunks]
    tw = eng.terrain_world(77)
    want_eng = [T.chunk(tw.world, x, z)[0] for x, z in chunks]

    def check_kinds(G):
        for lo in range(0, len(chunks), G.max_chunks):
            part = slice(lo, lo + G.max_chunks)
            dens, veins, _, _ = G.batch(cx[part], cz[part])
            for i, (od, ov, _) in enumerate(want_kinds[part]):
                assert same_bits(dens[i], od) and np.array_equal(veins[i], ov), (G.max_chunks, lo + i)

    def check_eng(G):
        for lo in range(0, len(chunks), G.max_chunks):
            part = slice(lo, lo + G.max_chunks)
            dens, _, _, _ = G.batch(cx[part], cz[part])
            for i, od in enumerate(want_eng[part]):
                assert same_bits(dens[i], od), (G.max_chunks, lo + i)

    for k, mc in enumerate((2, 16, 1, 8)):
        print("terrain: kit reuse, max_chunks %d" % mc, flush=True)
        A = mcgpu.TerrainWorld(W, max_chunks=mc)
        B = mcgpu.TerrainWorld(tw.world, max_chunks=17 - mc)
        (check_kinds(A), check_eng(B)) if k % 2 == 0 else (check_eng(B), check_kinds(A))
        (A.close(), B.close()) if k % 2 == 0 else (B.close(), A.close())
    tw.close()
