"""Chunk-mode biomes (libmcgen world_chunk_biomes, the Java oracle's chunk grids) in the oracle.

* the Java oracle's 26.3 chunk-mode grid hashes (tests/vectors/26.3/biome_grid_hashes.tsv "chunk" rows:
  climate through the chunk / volume path, search history carried chunk by chunk) are reproduced;
* whole chunks through the volume path with a fresh fill-order history per chunk (or_chunk_biomes) equal
  the plugin restatement's point-path climate with the backward tie chain (or_mcg_cell): the volume order
  sums each octave's gradient terms (x + z) + y, the point order (x + y) + z, and every gradient vector
  has a zero component, so both round alike.
"""
import os

import numpy as np
import pytest

import mcgprog
import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vectors")


def test_chunk_grid_matches_java():
    rows = [l.rstrip("\n").split("\t") for l in open(os.path.join(GOLD, "26.3", "biome_grid_hashes.tsv"))
            if not l.startswith("#")]
    rows = [r for r in rows if r[2] == "chunk"]
    assert len(rows) == 6
    o = oracle.Oracle("26.3", "overworld")
    for dim, seed, mode, qx0, qz0, n, qy, h in rows:
        out = o.grid(int(seed), int(qx0), int(qz0), int(n), int(n), int(qy), semantics=2)
        assert "%016x" % oracle.fnv1a_names(out.reshape(-1)) == h, seed


@pytest.mark.parametrize("version,dim", [("26.3", "overworld"), ("26.3", "nether"), ("26.1", "overworld"),
                                         ("26.3", "end"), ("26.1", "end")])
def test_chunks_equal_chain_formulation(version, dim):
    seed = 77
    V = mcgprog.vanilla(version, dim, seed)
    rng = np.random.default_rng(21)
    chunks = rng.integers(-30_000, 30_000, (3, 2)).astype(np.int32)
    if dim == "end":
        chunks[0] = (3, -2)                       # inside the central island
    got = V.oracle.chunk_biomes(seed, chunks)
    qh, qmin = got.shape[1], V.world.min_qy
    for c, (cx, cz) in enumerate(chunks):
        for qy in range(qh):
            for qz in range(4):
                for qx in range(4):
                    b, unres, _ = V.ref.cell(int(cx) * 4 + qx, qmin + qy, int(cz) * 4 + qz)
                    assert not unres and b == got[c, qy, qz, qx], (c, qx, qy, qz)
