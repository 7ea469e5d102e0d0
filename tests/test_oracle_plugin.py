"""Pins of oracle/mcgplugin.c and the plugin-form world builder (or_mcg_world), the checker of the
mcgpu_* entry points (libmcgen GPU plugin ABI).

* the builder's program reproduces the oracle's climate (itself pinned by the reference's climate
  vectors) bit for bit, every version and dimension;
* the plugin-form world searched in row-major order with search history reproduces the Java oracle's
  grid hashes (tests/vectors/<V>/biome_grid_hashes.tsv, 26.1 and 26.2 "point" rows): program, tree
  layout and candidate search together against Java;
* a cell that is not tied gets the engine's biome (no candidate);
* the backward tie chain (world_biome_cell) equals one sequential pass over the chunk in fill order
  (two formulations of the same history), on vanilla chunks with ties and on worlds whose leaves
  overlap so that ties are everywhere.
"""
import os

import numpy as np
import pytest

import mcgprog
import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vectors")


@pytest.mark.parametrize("version", ["26.1", "26.3"])
@pytest.mark.parametrize("dim", ["overworld", "nether", "end"])
def test_program_reproduces_oracle_climate(version, dim):
    V = mcgprog.vanilla(version, dim, 12345)
    o = V.oracle
    rng = np.random.default_rng(7)
    for _ in range(200):
        bx, bz = (int(v) for v in rng.integers(-3_000_000, 3_000_000, 2))
        by = int(rng.integers(-64, 320))
        got = V.ref.climate(bx, by, bz)
        if dim == "overworld":
            want = o.climate_raw(V.seed, bx, by, bz)
            assert [np.float32(g) for g in got] == [np.float32(w) for w in want], (bx, by, bz)
        elif dim == "end":
            sx, sz = int(np.int32(bx) // 8 if bx >= 0 else -((-bx) // 8)), int(bz // 8 if bz >= 0 else -((-bz) // 8))
            sq = np.int32(np.uint32((sx * sx + sz * sz) & 0xFFFFFFFF).view(np.int32))
            if version != "26.3" and sq < 0:
                # the 26.1/26.2 centre term's int square wraps negative: Java's clamp and max keep the NaN
                # (libmcgen, df_old.c:449-465) where the engine's clamp gives 80 (engine/mc_end.h)
                assert np.isnan(got[3]), (bx, bz)
            else:
                assert np.float32(got[3]) == np.float32(o.end_erosion(V.seed, bx, bz)), (bx, bz)
        else:
            q = np.array([[bx >> 2, by >> 2, bz >> 2]], np.int32)
            _, tg = o.points(np.array([V.seed], np.int64), q, want_tg=True)
            c = V.ref.climate((bx >> 2) * 4, (by >> 2) * 4, (bz >> 2) * 4)
            assert [mcgprog.quantize(c[0]), mcgprog.quantize(c[1])] == [int(tg[0][0][0]), int(tg[0][0][1])]
            assert c[2:] == [0.0, 0.0, 0.0, 0.0]


def _rows(version):
    out = []
    for line in open(os.path.join(GOLD, version, "biome_grid_hashes.tsv")):
        if line.startswith("#"):
            continue
        dim, seed, mode, qx0, qz0, n, qy, h = line.rstrip("\n").split("\t")
        if mode == "point" and dim != "the_end" and int(seed) in (0, -1):
            out.append((dim, int(seed), int(qx0), int(qz0), int(n), int(qy), h))
    return out


@pytest.mark.parametrize("version", ["26.1", "26.2"])
def test_history_grid_matches_java(version):
    rows = _rows(version)
    assert len(rows) == 4
    for dim, seed, qx0, qz0, n, qy, h in rows:
        V = mcgprog.vanilla(version, dim, seed)
        out, ok = V.ref.quart_history(qx0, qz0, n, n, qy)
        assert ok
        assert "%016x" % oracle.fnv1a_names(out) == h, (version, dim, seed)


@pytest.mark.parametrize("version", ["26.3", "26.1"])
def test_untied_cell_is_engine_biome(version):
    V = mcgprog.vanilla(version, "overworld", 99)
    rng = np.random.default_rng(3)
    cells = []
    for _ in range(400):
        qx, qz = (int(v) for v in rng.integers(-200_000, 200_000, 2))
        qy = int(rng.integers(-30, 110))
        cells.append((qx, qy, qz))
    qyc = [min(max(q[1], -16), 79) for q in cells]
    pts = np.array([[q[0], y, q[2]] for q, y in zip(cells, qyc)], np.int32)
    eng = V.oracle.points(np.array([99], np.int64), pts)[0]
    for (qx, qy, qz), e in zip(cells, eng):
        b, unres, chain = V.ref.cell(qx, qy, qz)
        assert unres == 0
        if chain == 0:
            assert b == e, (qx, qy, qz)


def _chain_vs_seq(ref, cx, cz, qh, min_qy, upto=None):
    seq, valid = ref.chunk_seq(cx, cz, qh)
    n = 0
    for qy in range(upto or qh):
        for qz in range(4):
            for qx in range(4):
                b, unres, chain = ref.cell(cx * 4 + qx, min_qy + qy, cz * 4 + qz)
                if valid[qy, qz, qx]:
                    assert b == seq[qy, qz, qx], (cx, cz, qx, qy, qz, chain)
                    n += chain > 0
    return n


def test_chain_equals_fill_order_history_vanilla():
    V = mcgprog.vanilla("26.3", "overworld", 5)
    rng = np.random.default_rng(11)
    chunks, tied = set(), 0
    while len(chunks) < 3:
        qx, qz = (int(v) for v in rng.integers(-100_000, 100_000, 2))
        qy = int(rng.integers(-16, 80))
        if V.ref.cell(qx, qy, qz)[2] > 0:
            chunks.add((qx >> 2, qz >> 2))
    for cx, cz in chunks:
        tied += _chain_vs_seq(V.ref, cx, cz, 96, -16)
    assert tied >= 3


def overlap_world(version, seed, shift):
    """Vanilla program; a flat tree of the vanilla leaves followed by a copy of every third leaf shifted
    by `shift` along temperature, with another biome: where a copy overlaps the original the cell is tied
    and its biome depends on the cells before it in fill order."""
    V = mcgprog.vanilla(version, "overworld", seed)
    nodes = mcgprog.tree_nodes(V.world)
    leaves = [(np.array(n.lo[:6]), np.array(n.hi[:6]), n.biome) for n in nodes if n.count == 0 and n.lo[6] == 0 and n.hi[6] == 0]
    extra = []
    for i, (lo, hi, b) in enumerate(leaves):
        if i % 3 == 0:
            d = np.zeros(6, np.int64)
            d[0] = shift
            extra.append((lo + d, hi + d, (b + 1) % 60))
    T = mcgprog.flat_tree(leaves + extra)
    return V, mcgprog.world(V.prog, V.world, tree=T, keep=V)


def overlap_world_snapped(version, seed):
    """Like overlap_world, but every third leaf's copy widens its temperature interval down to the next
    lower leaf bound already in the tree, so the copies add no breakpoint (the plugin keeps its cell
    table) while overlapping their neighbours (ties at distance 0)."""
    V = mcgprog.vanilla(version, "overworld", seed)
    nodes = mcgprog.tree_nodes(V.world)
    leaves = [(np.array(n.lo[:6]), np.array(n.hi[:6]), n.biome) for n in nodes if n.count == 0 and n.lo[6] == 0 and n.hi[6] == 0]
    bps = sorted({int(l[0][0]) for l in leaves} | {int(l[1][0]) for l in leaves})
    extra = []
    for i, (lo, hi, b) in enumerate(leaves):
        below = [v for v in bps if v < lo[0]]
        if i % 3 == 0 and below:
            lo2 = lo.copy()
            lo2[0] = below[-1]
            extra.append((lo2, hi.copy(), (b + 1) % 60))
    T = mcgprog.flat_tree(leaves + extra)
    return V, mcgprog.world(V.prog, V.world, tree=T, keep=V)


def test_chain_equals_fill_order_history_overlapping():
    import ctypes as C
    # the snapped world ties almost everywhere: its chains run back to the chunk's first cell, so only the
    # lowest two sections are compared there
    for make, upto in ((lambda: overlap_world("26.3", 8, 1500), None), (lambda: overlap_world_snapped("26.1", 8), 8)):
        V, W = make()
        ref = oracle.McgWorldRef(C.addressof(W), keep=W)
        rng = np.random.default_rng(5)
        tied = 0
        for _ in range(3):
            cx, cz = (int(v) for v in rng.integers(-20_000, 20_000, 2))
            tied += _chain_vs_seq(ref, cx, cz, 96, -16, upto)
        assert tied > 20


def test_grid_equals_points():
    V = mcgprog.vanilla("26.1", "overworld", 42)
    g, bad = V.ref.grid(-301, 517, 37, 29, 3, 70)
    xs = np.arange(37) * 3 - 301
    zs = np.arange(29) * 3 + 517
    xyz = np.array([[x, 70, z] for z in zs for x in xs], np.int32)
    p, bad2 = V.ref.points(xyz)
    assert bad == bad2 == 0
    assert (g.reshape(-1) == p).all()
