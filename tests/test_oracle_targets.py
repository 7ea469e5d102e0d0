"""The oracle's target search (or_targets, the restated engine/mc_biomes.h:31-62) pinned by brute force:
its biome is among the least-distance leaves of all leaves (or_tie_mask scans every leaf), on targets
that sit on, beside and between the parameter breakpoints and far outside them; and on the reference's
Java-oracle climate vectors it gives the recorded biome wherever that minimum is unique."""
import numpy as np
import pytest

import golden_io as G
import oracle as O


def breakpoints(orc):
    T = orc.tree_serialize()
    leaves = T[T[:, 0] == 1]
    return [np.unique(np.concatenate([leaves[:, 1 + 2 * k], leaves[:, 2 + 2 * k]])) for k in range(6)]


def adversarial_targets(orc, rng, n):
    """Targets whose coordinates sit on, next to and between breakpoints, plus wide random values."""
    cols = []
    for b in breakpoints(orc):
        mids = (b[:-1] + b[1:]) // 2 if b.size > 1 else b
        pool = np.unique(np.concatenate([b, b - 1, b + 1, mids, [0, 32767, -32767, 32768, -32768, 40000, -40000]]))
        pick = pool[rng.integers(0, pool.size, n)]
        wide = rng.integers(-15000, 15001, n)
        far = rng.integers(-(1 << 20), (1 << 20) + 1, n)
        u = rng.random(n)
        cols.append(np.where(u < 0.7, pick, np.where(u < 0.95, wide, far)))
    return np.stack(cols, 1).astype(np.int32)


@pytest.mark.parametrize("version", G.VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether"])
def test_targets_are_least_distance(version, dim):
    orc = O.Oracle(version, dim)
    tg = adversarial_targets(orc, np.random.default_rng(5 + G.VERSIONS.index(version)), 1500)
    if dim == "the_nether":
        tg[:, 2:] = 0
    got = orc.targets(tg)
    for t, b in zip(tg, got):
        cnt, m = orc.tie_mask(t)
        assert cnt >= 1 and (m[int(b) >> 6] >> (int(b) & 63)) & 1, (t.tolist(), int(b))


@pytest.mark.parametrize("version", G.VERSIONS)
def test_targets_match_java_vectors(version):
    rows = G.climate_rows(version, "overworld")
    names = O.biome_names()
    orc = O.Oracle(version, "overworld")
    tg = np.array([r[4] for r in rows], np.int32)
    got = orc.targets(tg)
    bad = 0
    for r, t, b in zip(rows, tg, got):
        if names[b] != r[5]:
            cnt, _ = orc.tie_mask(t)
            bad += cnt < 2
    assert bad == 0
