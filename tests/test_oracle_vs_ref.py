"""The restatement against the reference's own CPU engine (engine/cuda/mcg_host.c compiled from its
sources into oracle/_ref by `make -C oracle ref`): identical biomes and climate targets on the same
inputs, including far coordinates, every dimension and version.  Skipped where oracle/_ref is not built."""
import numpy as np
import pytest

import golden_io as G
import oracle as O

pytestmark = pytest.mark.skipif(not O.RefEngine.available(), reason="oracle/_ref not built (make -C oracle ref)")


def _inputs(rng, ns, np_, span):
    seeds = rng.integers(-2**63, 2**63 - 1, ns, dtype=np.int64)
    seeds[:5] = [0, 1, -1, 2**63 - 1, -2**63]
    pts = np.stack([rng.integers(-span, span, np_), rng.integers(-16, 80, np_), rng.integers(-span, span, np_)], 1)
    return seeds, pts.astype(np.int32)


@pytest.mark.parametrize("version", G.VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether", "the_end"])
def test_points_identical(version, dim):
    rng = np.random.default_rng(hash((version, dim)) & 0xFFFF)
    ns, np_ = (24, 64) if dim != "the_end" else (8, 24)
    for span in (2**10, 2**20, 2**22):
        seeds, pts = _inputs(rng, ns, np_, span >> 2)
        a, ta = O.Oracle(version, dim).points(seeds, pts, want_tg=True)
        ref = O.RefEngine(version, dim)
        b, tb = ref.points(seeds, pts, want_tg=True)
        ref.close()
        assert np.array_equal(a, b)
        assert np.array_equal(ta, tb)


@pytest.mark.parametrize("preset", ["amplified", "large_biomes"])
def test_presets_identical(preset):
    rng = np.random.default_rng(3)
    seeds, pts = _inputs(rng, 16, 64, 2**18)
    for version in G.VERSIONS:
        a, ta = O.Oracle(version, "overworld", preset).points(seeds, pts, want_tg=True)
        ref = O.RefEngine(version, "overworld", preset)
        b, tb = ref.points(seeds, pts, want_tg=True)
        ref.close()
        assert np.array_equal(a, b) and np.array_equal(ta, tb)
