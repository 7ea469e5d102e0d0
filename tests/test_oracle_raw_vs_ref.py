"""Raw climate floats and NormalNoise values of the oracle against the reference's own header-only engine
(oracle/ref_shim.c over engine/*.h, built from the reference sources into oracle/_ref), bit for bit.
Quantized targets cannot pin PerlinNoise.wrap (engine/mc_noise.h:77, 86-89): wrap removes multiples of
2^25, a multiple of the lattice period 256, so it changes only the rounding of the fractional part, far
below the 1e-4 quantization step.  Raw values show it: positions are drawn past 2^24 / f so that wrap's
outer branch runs in every octave."""
import struct
import zlib

import numpy as np
import pytest

import golden_io as G
import oracle as O

ALL_VERSIONS = G.VERSIONS + ("26.4-snapshot-2",)

pytestmark = pytest.mark.skipif(not O.RefShim.available(), reason="oracle/_ref not built (make -C oracle ref)")


def _blocks(rng, n, lo, hi):
    mag = 2.0 ** rng.uniform(lo, hi, (n, 2))
    xz = (np.minimum(mag, 2**30 - 1).astype(np.int64) * rng.choice([-1, 1], (n, 2)))
    return np.stack([xz[:, 0], rng.integers(-64, 320, n), xz[:, 1]], 1).astype(np.int32)


@pytest.mark.parametrize("version", ALL_VERSIONS)
@pytest.mark.parametrize("preset", ["overworld", "amplified", "large_biomes"])
def test_raw_climate_bits_identical(version, preset):
    rng = np.random.default_rng(zlib.crc32(("raw/%s/%s" % (version, preset)).encode()))
    shim = O.RefShim(version, preset)
    orc = O.Oracle(version, "overworld", "normal" if preset == "overworld" else preset)
    for seed in (0, -1, int(rng.integers(-2**63, 2**63 - 1, dtype=np.int64))):
        for lo, hi in ((0, 22), (24, 30)):
            xyz = _blocks(rng, 48, lo, hi)
            want = shim.climate_raw(seed, xyz)
            got = np.array([orc.climate_raw(seed, *map(int, p)) for p in xyz], np.float32)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (seed, lo, hi)
    shim.close()


NOISES = ["offset", "temperature", "vegetation", "continentalness", "erosion", "ridge"]


@pytest.mark.parametrize("version", ALL_VERSIONS)
def test_noise_values_bits_identical(version):
    """NormalNoise at double positions of every magnitude up to 2^36 (octave inputs x*f past 2^24 for
    every frequency of these noises), Overworld and legacy Nether noises."""
    rng = np.random.default_rng(zlib.crc32(("noise/%s" % version).encode()))
    shim = O.RefShim(version)
    orc = O.Oracle(version, "overworld")
    xyz = (2.0 ** rng.uniform(-4, 36, (64, 3))) * rng.choice([-1.0, 1.0], (64, 3)) + rng.random((64, 3))
    xyz[:4] = [[0, 0, 0], [2**24, -2**24, 0.5], [-(2**24 - 0.5), 3.25, 2**25 + 0.75], [1e10, -1e10, 123.456]]
    for name in NOISES + ["nether/temperature", "nether/vegetation"]:
        for seed in (0, 12345, -2**63):
            want = shim.noise(name, seed, xyz)
            got = np.array([orc.noise(name, seed, *p) for p in xyz])
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (name, seed)
    shim.close()
