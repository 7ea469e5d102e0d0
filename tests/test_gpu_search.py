"""Seed search on the device (mcb_search) against the reference's known answers and the oracle.

Known answers: crack/src/gpu_biomes/results/example_obs_26.3_overworld_seed*.txt (8 observations of the
Java oracle for seeds 8675309 and -7048155917072976836; the reference's full int32 scan finds exactly
8675309, research_full32_ow_26.3.txt).  Oracle cross-checks: candidate lists around a seed whose own
biomes make the observations, every dimension and version, quart and block (BiomeManager) points,
ties, the pruned Overworld kernel against the slot path and against no pruning."""
import os

import numpy as np
import pytest

import golden_io as G
import oracle as O

pytestmark = pytest.mark.gpu
mcbiome = pytest.importorskip("mcbiome")
NAMES = O.biome_names()
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", autouse=True)
def _device():
    if mcbiome.device_count() < 1:
        pytest.fail("no HIP device visible to libmcbiome_hip.so")


def _kat(name):
    obs = []
    for line in open(os.path.join(HERE, "golden", "search", name)):
        if line.startswith("#") or not line.strip():
            continue
        x, y, z, b = line.split()
        obs.append(mcbiome.make_obs(int(x), int(y), int(z), [NAMES.index(b.replace("minecraft:", ""))]))
    return obs


@pytest.mark.parametrize("seed,name", [(8675309, "example_obs_26.3_overworld_seed8675309.txt"),
                                       (-7048155917072976836, "example_obs_26.3_overworld_seed-7048155917072976836.txt")])
def test_known_answer_window(seed, name):
    """A 2^20-candidate window around the answer finds exactly the answer, and every candidate is counted."""
    obs = _kat(name)
    eng = mcbiome.Engine("26.3", "overworld")
    n = 1 << 20
    found, reached, ms = eng.search(obs, base=seed - (n >> 1), n=n)
    print("kat", seed, "reached", reached.tolist(), "ms %.1f" % ms, "candidates/s %.3g" % (n / ms * 1e3), flush=True)
    assert found.tolist() == [seed]
    assert int(reached.sum()) == n and int(reached[-1]) == 1


def _observations(orc, seed, rng, n, block=False, extra=0, span=3000, qy=(-10, 70)):
    obs_ref, obs_dev = [], []
    for i in range(n):
        x, z = int(rng.integers(-span, span)), int(rng.integers(-span, span))
        y = int(rng.integers(*qy))
        blk = block and i % 2 == 1
        if blk:
            q = O.zoom(O.hashed_seed(seed), x * 4 + 1, y * 4 + 2, z * 4 + 3)
            b = int(orc.points([seed], np.array([q], np.int32))[0, 0])
            pt = (x * 4 + 1, y * 4 + 2, z * 4 + 3)
        else:
            b = int(orc.points([seed], np.array([[x, y, z]], np.int32))[0, 0])
            pt = (x, y, z)
        ids = {b} | {int(v) for v in rng.integers(0, len(NAMES), extra)}
        obs_ref.append(pt + (blk, ids))
        obs_dev.append(mcbiome.make_obs(*pt, sorted(ids), block=blk))
    return obs_ref, obs_dev


def _candidates(rng, seed, n):
    c = [seed, seed + 1, seed - 1, -seed, seed ^ (1 << 48)] + [int(v) for v in rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)]
    return np.array(c, np.int64)


@pytest.mark.parametrize("version", G.VERSIONS)
def test_overworld_search_matches_oracle(version):
    rng = np.random.default_rng(11 + G.VERSIONS.index(version))
    orc = O.Oracle(version, "overworld")
    eng = mcbiome.Engine(version, "overworld")
    seed = int(rng.integers(-2**62, 2**62))
    obs_ref, obs_dev = _observations(orc, seed, rng, 6, block=True, extra=2)
    cands = _candidates(rng, seed, 3000)
    want_f, want_r = orc.search(cands, obs_ref)
    print("overworld", version, "reached", want_r.tolist(), flush=True)
    for flags in (0, mcbiome.SEARCH_NOPRUNE, mcbiome.SEARCH_SLOTS):
        f, r, _ = eng.search(obs_dev, seeds=cands, flags=flags)
        assert r.tolist() == want_r.tolist(), flags
        assert f.tolist() == want_f.tolist(), flags
    assert seed in want_f.tolist()


@pytest.mark.parametrize("version", G.VERSIONS)
def test_overworld_ties_match_oracle(version):
    """With ties an observation accepts any least-distance leaf's biome: sets drawn from the leaves
    tied at the true seed's points where there are ties, plus ordinary points."""
    rng = np.random.default_rng(31 + G.VERSIONS.index(version))
    orc = O.Oracle(version, "overworld")
    eng = mcbiome.Engine(version, "overworld")
    seed = int(rng.integers(-2**62, 2**62))
    obs_ref, obs_dev = _observations(orc, seed, rng, 5, extra=1)
    cands = _candidates(rng, seed, 1500)
    want_f, want_r = orc.search(cands, obs_ref, ties=True)
    for flags in (mcbiome.SEARCH_TIES, mcbiome.SEARCH_TIES | mcbiome.SEARCH_SLOTS):
        f, r, _ = eng.search(obs_dev, seeds=cands, flags=flags)
        assert r.tolist() == want_r.tolist() and f.tolist() == want_f.tolist(), flags


@pytest.mark.parametrize("dim", ["the_nether", "the_end"])
def test_nether_end_search_matches_oracle(dim):
    for version in G.VERSIONS:
        rng = np.random.default_rng(51 + G.VERSIONS.index(version))
        orc = O.Oracle(version, dim)
        eng = mcbiome.Engine(version, dim)
        seed = int(rng.integers(-2**62, 2**62))
        span = 3000 if dim == "the_nether" else 1200
        obs_ref, obs_dev = _observations(orc, seed, rng, 5, block=True, extra=1, span=span)
        cands = _candidates(rng, seed, 1500)
        want_f, want_r = orc.search(cands, obs_ref)
        f, r, _ = eng.search(obs_dev, seeds=cands)
        print(dim, version, "reached", want_r.tolist(), flush=True)
        assert r.tolist() == want_r.tolist() and f.tolist() == want_f.tolist()
        if dim == "the_end":
            # the narrowed cell window decides every observation as the full window does
            for o in obs_dev:
                ids = [b for b in range(len(NAMES)) if (o.mask[b >> 6] >> (b & 63)) & 1]
                o.flags = (o.flags & ~0xFF) | eng.end_obs_radius(ids)
            f2, r2, _ = eng.search(obs_dev, seeds=cands)
            assert r2.tolist() == want_r.tolist() and f2.tolist() == want_f.tolist()


def test_search_sources_and_limits():
    eng = mcbiome.Engine("26.3", "overworld")
    obs = _kat("example_obs_26.3_overworld_seed8675309.txt")
    # structural source: 65536 candidates base | i << 48
    f, r, _ = eng.search(obs[:2], base=8675309, structural=True)
    assert int(r.sum()) == 65536
    structural = [(8675309 | (i << 48)) - ((1 << 64) if i >= 32768 else 0) for i in range(65536)]
    f1, r1, _ = eng.search(obs[:2], seeds=structural)
    assert r.tolist() == r1.tolist() and f.tolist() == sorted(f1.tolist(), key=lambda v: v & ((1 << 64) - 1))
    # max_found truncates the list, not the count
    f2, r2, _ = eng.search(obs[:1], base=0, n=4096, max_found=3)
    assert len(f2) == min(3, int(r2[-1]))
    with pytest.raises(mcbiome.EngineError):
        eng.search([], base=0, n=10)
    with pytest.raises(mcbiome.EngineError):
        eng.search(obs * 9, base=0, n=10)
