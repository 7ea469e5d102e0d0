"""The restatement against the reference's own CPU engine (engine/cuda/mcg_host.c compiled from its
sources into oracle/_ref by `make -C oracle ref`): identical biomes and climate targets on the same
inputs, including far coordinates, every dimension and version.  Skipped where oracle/_ref is not built."""
import zlib

import numpy as np
import pytest

import golden_io as G
import oracle as O

ALL_VERSIONS = G.VERSIONS + ("26.4-snapshot-2",)

pytestmark = pytest.mark.skipif(not O.RefEngine.available(), reason="oracle/_ref not built (make -C oracle ref)")


def _inputs(rng, ns, np_, span):
    seeds = rng.integers(-2**63, 2**63 - 1, ns, dtype=np.int64)
    seeds[:5] = [0, 1, -1, 2**63 - 1, -2**63]
    pts = np.stack([rng.integers(-span, span, np_), rng.integers(-16, 80, np_), rng.integers(-span, span, np_)], 1)
    return seeds, pts.astype(np.int32)


@pytest.mark.parametrize("version", ALL_VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether", "the_end"])
def test_points_identical(version, dim):
    rng = np.random.default_rng(zlib.crc32(("%s/%s" % (version, dim)).encode()))
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
    for version in ALL_VERSIONS:
        a, ta = O.Oracle(version, "overworld", preset).points(seeds, pts, want_tg=True)
        ref = O.RefEngine(version, "overworld", preset)
        b, tb = ref.points(seeds, pts, want_tg=True)
        ref.close()
        assert np.array_equal(a, b) and np.array_equal(ta, tb)


@pytest.mark.parametrize("version", ALL_VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether", "the_end"])
def test_far_coordinates_identical(version, dim):
    """Quart coordinates of magnitude 2^24 .. 2^28 (blocks up to 2^30, inside int32): the octave inputs
    x*f pass nextDown(2^24) in the shift and climate noises, so PerlinNoise.wrap's outer branch
    (engine/mc_noise.h:77, 86-89) decides the result.  The restatement equals the reference engine."""
    rng = np.random.default_rng(zlib.crc32(("far/%s/%s" % (version, dim)).encode()))
    ns, np_ = (8, 96) if dim != "the_end" else (4, 32)
    seeds = rng.integers(-2**63, 2**63 - 1, ns, dtype=np.int64)
    seeds[:3] = [0, 1, -1]
    mag = 2.0 ** rng.uniform(24, 28, (np_, 2))
    sign = rng.choice([-1, 1], (np_, 2))
    xz = np.minimum(mag, 2**28 - 1).astype(np.int64) * sign
    pts = np.stack([xz[:, 0], rng.integers(-16, 80, np_), xz[:, 1]], 1).astype(np.int32)
    pts[:4, 0] = [2**28 - 1, -2**28, 1 << 24, -(1 << 24)]
    a, ta = O.Oracle(version, dim).points(seeds, pts, want_tg=True)
    ref = O.RefEngine(version, dim)
    b, tb = ref.points(seeds, pts, want_tg=True)
    ref.close()
    assert np.array_equal(ta, tb)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("version", ALL_VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether"])
def test_tie_mask_identical(version, dim):
    """or_tie_mask against the reference's mcg_ref_tie_mask (engine/cuda/mcg_host.c:85-96): targets of
    real samples, the same with coordinates moved onto nearby parameter breakpoints (ties), and
    far-off targets of int32 magnitude."""
    rng = np.random.default_rng(zlib.crc32(("tie/%s/%s" % (version, dim)).encode()))
    seeds, pts = _inputs(rng, 6, 48, 2**16)
    _, tg = O.Oracle(version, dim).points(seeds, pts, want_tg=True)
    tg = tg.reshape(-1, 6)
    snapped = (np.round(tg / 500.0) * 500).astype(np.int32)
    far = rng.integers(-2**30, 2**30, (32, 6)).astype(np.int32)
    orc, ref = O.Oracle(version, dim), O.RefEngine(version, dim)
    for t in np.concatenate([tg, snapped, far]):
        cnt, m = orc.tie_mask(t)
        assert m == ref.tie_mask(t), t
    ref.close()
