"""The terrain restatement (oracle/mcgterrain.c) against libmcgen's own density-function pipeline
(oracle/_ref/libmcgen_ref.so, compiled from libmcgen/src by `make -C oracle ref`): the same exported
program, the same volumes, bit for bit; the cache cells a chunk leaves behind; the ore random.  Skipped
where the reference build is absent."""
import numpy as np
import pytest

import terrain as T
import terrain_worlds as TW

pytestmark = pytest.mark.skipif(not T.RefTerrain.available(), reason="libmcgen reference build absent (make -C oracle ref)")


def same_bits(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return a.shape == b.shape and bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))))


VOLS = [(16, 48, 16, 32, -64, -48, 1, 1, 1),          # a block slab of a chunk
        (5, 7, 5, -16, -64, 32, 4, 8, 4),              # cell corners
        (7, 13, 3, -1000003, 100, 999, 1, 1, 1),       # off the cell grid, odd sizes, far out
        (3, 5, 4, 8, -64, 8, 4, 8, 4),
        (1, 1, 1, 5, 70, -9, 1, 1, 1)]                 # one block


@pytest.mark.parametrize("world", ["kinds", "overworld_like"])
def test_restatement_equals_libmcgen_volumes(world):
    if world == "kinds":
        noises, funcs, fd, _ = TW.kinds_world()
    else:
        noises, funcs, fd = TW.overworld_like()
    for seed in (0, -4172144997902289642):
        R = T.RefTerrain(seed, noises, funcs)
        r = R.compile(fd)
        P = R.export([r])
        node = P.contents.root[0]
        for vol in VOLS:
            ref = R.volume(r, vol)
            assert same_bits(T.volume(P, node, vol), ref), (world, seed, vol)
            # values do not depend on the context's caches
            assert same_bits(R.volume(r, vol, caches=False), ref)


def test_world_covers_every_volume_kind():
    """The kinds world's exported program holds every node kind the volume executor runs."""
    import mcgpu
    noises, funcs, fd, _ = TW.kinds_world()
    R = T.RefTerrain(1, noises, funcs)
    P = R.export([R.compile(fd)]).contents
    kinds = {P.nodes[i].k for i in range(P.nnodes)}
    # log, pow, round and find_top_surface are refused by the volume executor (as by the reference's)
    covered = set(mcgpu.K.values()) - {mcgpu.K[k] for k in ("LOG", "ROUND_INT", "ROUND", "POW_CB", "POW_CE", "POW", "FTS")}
    assert covered <= kinds, sorted(covered - kinds)


def test_chunk_density_cells_and_ore_random():
    noises, funcs, fd, pre = TW.kinds_world()
    dv, rv, gv = TW.vein_roots()
    R = T.RefTerrain(-4172144997902289642, noises, funcs)
    r, p, d, ri, gp = (R.compile(x) for x in (fd, pre, dv, rv, gv))
    cids = sorted(set(R.cache_ids(p) + R.cache_ids(r)))
    W = T.terrain_world(R, r, veins=[(d, ri, gp, 11, 12, 13, 0.02)], pre=p, cache_ids=cids)
    for cx, cz in [(0, 0), (-3, 5), (70001, -65535)]:
        dens, veins, cells = T.chunk(W, cx, cz)
        rd, rcells = T.ref_chunk(R, W, r, cx, cz, pre=p)
        assert same_bits(dens, rd)
        assert [c[:3] for c in cells] == [c[:3] for c in rcells]
        assert all(same_bits(a[3], b[3]) for a, b in zip(cells, rcells))
        assert any(c[0] and c[1] for c in cells) and 0 < int((veins != 0xFFFF).sum()) < len(veins)
        # the vein patch: one of the vein's three block states wherever it is set
        assert set(np.unique(veins)) <= {11, 12, 13, 0xFFFF}
    # Mth.getSeed positional Xoroshiro floats: restatement against libmcgen's pos_at / nextFloat
    xyz = np.array([[0, 0, 0], [5, -60, 7], [-1000, 100, 33333], [123456, 5, -654321], [-2**25, 319, 2**25 - 1]], np.int32)
    lo, hi = R.ore()
    ref = R.ore_floats(xyz, 4)
    mine = np.array([T.ore_floats(lo, hi, *map(int, q), 4) for q in xyz])
    assert same_bits(mine, ref)


def test_interpolation_closed_forms():
    """Cell interpolation pinned by what the arithmetic fixes: cell corners reproduce the corner values
    exactly, a constant input stays constant, and a y gradient with exactly representable steps is
    reproduced exactly between corners."""
    R = T.RefTerrain(5, {}, {})
    grad = {"type": "gradient", "axis": "y", "from_coordinate": 0, "to_coordinate": 8, "from_value": 0.0, "to_value": 8.0}
    r = R.compile({"type": "interpolated", "cell_size_xz": 4, "cell_size_y": 8, "input": grad})
    c = R.compile({"type": "interpolated", "cell_size_xz": 4, "cell_size_y": 8, "input": 0.375})
    P = R.export([r, c]).contents
    vol = (16, 32, 16, -32, -16, 48, 1, 1, 1)
    out = T.volume(R.progs[False], P.root[0], vol).reshape(16, 16, 32)      # [iz][ix][iy]
    ys = np.arange(-16, 16, dtype=np.float32)
    assert np.array_equal(out, np.broadcast_to(np.clip(ys, 0, 8), out.shape))
    assert same_bits(out.ravel(), R.volume(r, vol))
    const = T.volume(R.progs[False], P.root[1], vol)
    assert np.all(const == np.float32(0.375))
