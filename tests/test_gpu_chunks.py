"""mcb_chunk_biomes (whole chunks as libmcgen stores them) against the oracle's or_chunk_biomes, which is
pinned by the Java oracle's 26.3 chunk-mode grids: every version and dimension, chunks far out, near the
origin and around the End's central island, and a batch large enough to use every workgroup."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
mcbiome = pytest.importorskip("mcbiome")


@pytest.fixture(scope="module", autouse=True)
def _device():
    if mcbiome.device_count() < 1:
        pytest.fail("no HIP device visible to libmcbiome_hip.so")


@pytest.mark.parametrize("version", ["26.1", "26.2", "26.3"])
@pytest.mark.parametrize("dim", ["overworld", "nether", "end"])
def test_chunks_match_oracle(version, dim):
    seed = -31337 if version != "26.2" else 2 ** 62 + 5
    rng = np.random.default_rng(len(version) * 7 + len(dim))
    chunks = np.concatenate([rng.integers(-200_000, 200_000, (40, 2)), rng.integers(-8, 8, (8, 2)),
                             [[-2_000_000, 1_500_000], [0, 0]]]).astype(np.int32)
    print("chunks", version, dim, len(chunks), flush=True)
    got = mcbiome.Engine(version, dim).chunk_biomes(seed, chunks)
    want = oracle.Oracle(version, dim).chunk_biomes(seed, chunks)
    assert got.shape == want.shape
    assert np.array_equal(got, want), int((got != want).sum())


def test_many_chunks():
    seed = 99
    rng = np.random.default_rng(3)
    chunks = rng.integers(-50_000, 50_000, (600, 2)).astype(np.int32)
    eng = mcbiome.Engine("26.3", "overworld")
    got = eng.chunk_biomes(seed, chunks)
    sub = rng.choice(len(chunks), 24, replace=False)
    want = oracle.Oracle("26.3", "overworld").chunk_biomes(seed, chunks[sub])
    assert np.array_equal(got[sub], want)
    assert np.array_equal(eng.chunk_biomes(seed, chunks[:5]), got[:5])
