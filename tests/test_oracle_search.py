"""The oracle's seed search (or_search, the checker of the device search): the reference's known answers
(observations of the Java oracle for two 26.3 seeds, crack/src/gpu_biomes/results/example_obs_*.txt) are
found in a window of candidates, and the C search equals the same semantics composed from the oracle's
points, zoom, hashed seed and brute-force tie masks (search_ref.py) on lists with block points and ties."""
import os

import numpy as np
import pytest

import oracle as O
import search_ref

NAMES = O.biome_names()
HERE = os.path.dirname(os.path.abspath(__file__))


def kat(name):
    obs = []
    for line in open(os.path.join(HERE, "golden", "search", name)):
        if line.startswith("#") or not line.strip():
            continue
        x, y, z, b = line.split()
        obs.append((int(x), int(y), int(z), False, {NAMES.index(b.replace("minecraft:", ""))}))
    return obs


@pytest.mark.parametrize("seed,name", [(8675309, "example_obs_26.3_overworld_seed8675309.txt"),
                                       (-7048155917072976836, "example_obs_26.3_overworld_seed-7048155917072976836.txt")])
def test_known_answers(seed, name):
    orc = O.Oracle("26.3", "overworld")
    cands = np.arange(seed - 2000, seed + 2000, dtype=np.int64)
    found, reached = orc.search(cands, kat(name))
    assert found.tolist() == [seed] and int(reached.sum()) == cands.size


@pytest.mark.parametrize("dim,ties", [("overworld", False), ("overworld", True), ("the_nether", False), ("the_end", False)])
def test_c_search_equals_composed(dim, ties):
    rng = np.random.default_rng(3)
    orc = O.Oracle("26.3", dim)
    seed = 424242
    obs = []
    for i in range(5):
        x, y, z = (int(v) for v in rng.integers(-900, 900, 3))
        blk = i % 2 == 1
        q = O.zoom(O.hashed_seed(seed), x, y, z) if blk else (x, y, z)
        b = int(orc.points([seed], np.array([q], np.int32))[0, 0])
        obs.append((x, y, z, blk, {b, int(rng.integers(0, len(NAMES)))}))
    cands = np.concatenate([[seed], rng.integers(-2**63, 2**63 - 1, 300, dtype=np.int64)]).astype(np.int64)
    f1, r1 = orc.search(cands, obs, ties=ties)
    f2, r2 = search_ref.search(orc, cands, obs, ties=ties)
    assert r1.tolist() == r2.tolist() and f1.tolist() == f2.tolist()
    assert seed in f1.tolist()
