"""The HIP engine's quart grids at the Java oracle's grid-hash inputs (biome_grid_hashes.tsv: 1024^2
quarts, End 512^2, at quart y 16 from (-512, -512)), one seed per version and dimension:
  * bit-identical to the oracle's grid under the engine's search semantics (no candidate); and
  * against the oracle's grid under the Java resolver's search history, whose FNV-1a hash is checked
    here against the row: every cell where the two differ is a fitness tie, i.e. both biomes are among
    the least-distance leaves of the cell's target (the reference's own tie filter, tests/diff_biomes.py:78-93)."""
import numpy as np
import pytest

import golden_io as G
import oracle as O

pytestmark = pytest.mark.gpu
mcbiome = pytest.importorskip("mcbiome")


@pytest.mark.parametrize("version", G.VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether", "the_end"])
def test_hip_grid_at_hash_inputs(version, dim):
    row = next(r for r in G.grid_hashes(version) if r[0] == dim and r[2] == "point")
    _, seed, _, qx0, qz0, n, qy, h = row
    print("grid hash", version, dim, seed, n, flush=True)
    got = mcbiome.Engine(version, dim).region(seed, qx0, qz0, n, n, [qy])[0]
    orc = O.Oracle(version, dim)
    assert np.array_equal(got, orc.grid(seed, qx0, qz0, n, n, qy, 0))
    hist = orc.grid(seed, qx0, qz0, n, n, qy, 1)
    assert O.fnv1a_names(hist) == h
    iz, ix = np.nonzero(got != hist)
    if dim == "the_end":
        assert iz.size == 0
        return
    pts = np.stack([qx0 + ix, np.full(ix.size, qy), qz0 + iz], 1).astype(np.int32)
    _, tg = orc.points([seed], pts, want_tg=True)
    for k in range(ix.size):
        cnt, m = orc.tie_mask(tg[0, k])
        a, b = int(got[iz[k], ix[k]]), int(hist[iz[k], ix[k]])
        assert cnt >= 2 and (m[a >> 6] >> (a & 63)) & 1 and (m[b >> 6] >> (b & 63)) & 1, (iz[k], ix[k])
    print("cells decided by the search history:", ix.size, "of", n * n, flush=True)
