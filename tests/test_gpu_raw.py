"""The engine-header entry points that return values rather than biomes, on the device:
  * mcb_climate_raw (mc_climate_overworld_raw, engine/mc_climate.h:113-164) against the Java density-
    function vectors df.tsv (continents, erosion, ridges, depth) and the oracle's raw floats, bit for bit,
    at block positions of every magnitude up to 2^30 (PerlinNoise.wrap's outer branch; the oracle's raw
    values there are pinned against the reference's own code on the CPU);
  * mcb_noise_values (normal_get_f / normal_get_d, engine/mc_noise.h:286-324) against the Java noise
    vectors noise.tsv (every climate and Nether noise row, arbitrary x, y, z) and the oracle, bit for bit."""
import zlib

import numpy as np
import pytest

import golden_io as G
import oracle as O

pytestmark = pytest.mark.gpu
mcbiome = pytest.importorskip("mcbiome")

DF_RAW = {"overworld/continents": 2, "overworld/erosion": 3, "overworld/ridges": 5, "overworld/depth": 4}
OW_NOISES = ["offset", "temperature", "vegetation", "continentalness", "erosion", "ridge"]
NE_NOISES = ["nether/temperature", "nether/vegetation"]


def _f32(v):
    return np.float32(v).view(np.uint32)


@pytest.mark.parametrize("version", G.VERSIONS)
def test_climate_raw_matches_df_vectors(version):
    eng = mcbiome.Engine(version, "overworld")
    n = 0
    for dim, seed, name, x, y, z, value, dbits, fbits in G.df_rows(version):
        if name not in DF_RAW:
            continue
        got = eng.climate_raw([seed], [[x, y, z]])[0, 0, DF_RAW[name]]
        want = np.uint32(fbits & 0xFFFFFFFF) if version == "26.3" else _f32(G.bits_to_double(dbits))
        assert got.view(np.uint32) == want, (seed, name, x, y, z)
        n += 1
    assert n == 160


@pytest.mark.parametrize("version", G.VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether"])
def test_climate_raw_matches_oracle(version, dim):
    rng = np.random.default_rng(zlib.crc32(("raw/%s/%s" % (version, dim)).encode()))
    eng = mcbiome.Engine(version, dim)
    orc = O.Oracle(version, dim)
    seeds = rng.integers(-2**63, 2**63 - 1, 6, dtype=np.int64)
    seeds[:3] = [0, -1, 2**63 - 1]
    mag = 2.0 ** rng.uniform(0, 30, (96, 2))
    xz = np.minimum(mag, 2**30 - 1).astype(np.int64) * rng.choice([-1, 1], (96, 2))
    xyz = np.stack([xz[:, 0], rng.integers(-64, 320, 96), xz[:, 1]], 1).astype(np.int32)
    got = eng.climate_raw(seeds, xyz)
    if dim == "overworld":
        for i, s in enumerate(seeds[:3]):
            want = np.array([orc.climate_raw(int(s), *map(int, p)) for p in xyz], np.float32)
            assert np.array_equal(got[i].view(np.uint32), want.view(np.uint32)), int(s)
    else:
        for i, s in enumerate(seeds[:3]):
            for j, p in enumerate(xyz[:32]):
                for k, name in enumerate(NE_NOISES):
                    v = orc.noise(name, int(s), p[0] * 0.25, p[1] * 0.0, p[2] * 0.25)
                    assert got[i, j, k].view(np.uint32) == _f32(v), (int(s), p.tolist(), name)
            assert not got[i, :, 2:].any()
    # at quart-aligned blocks, quantized (mc_quantize: float product, truncation) they are the points path's targets
    xq = xyz // 4
    raw = eng.climate_raw(seeds, xq * 4)
    _, tg = eng.points(seeds, xq, want_tg=True)
    q = (raw * np.float32(10000.0)).astype(np.int64)
    if dim != "overworld":
        q[:, :, 2:] = 0
    assert np.array_equal(q, tg)


@pytest.mark.parametrize("version", G.VERSIONS)
def test_noise_values_match_noise_vectors(version):
    engines = {"overworld": mcbiome.Engine(version, "overworld"), "the_nether": mcbiome.Engine(version, "the_nether")}
    n = 0
    for dim, seed, name, x, y, z, value, dbits, fbits in G.noise_rows(version):
        if dim == "overworld" and name in OW_NOISES:
            k = OW_NOISES.index(name)
        elif dim == "the_nether" and name in NE_NOISES:
            k = NE_NOISES.index(name)
        else:
            continue
        v = engines[dim].noise_values(k, [seed], [[x, y, z]])[0, 0]
        if version == "26.3":
            assert _f32(v) == np.uint32(fbits & 0xFFFFFFFF), (dim, seed, name, x, y, z)
        else:
            assert np.float64(v).view(np.int64) == np.int64(dbits), (dim, seed, name, x, y, z)
        n += 1
    assert n == 400


@pytest.mark.parametrize("version", G.VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether"])
def test_noise_values_match_oracle(version, dim):
    """Positions of every magnitude to 2^36 (wrap's outer branch in every octave), signed zeros."""
    rng = np.random.default_rng(zlib.crc32(("nv/%s/%s" % (version, dim)).encode()))
    eng = mcbiome.Engine(version, dim)
    orc = O.Oracle(version, dim)
    xyz = (2.0 ** rng.uniform(-4, 36, (80, 3))) * rng.choice([-1.0, 1.0], (80, 3)) + rng.random((80, 3))
    xyz[:4] = [[0.0, -0.0, 0.0], [2**24, -2**24, 0.5], [-(2**24 - 0.5), 3.25, 2**25 + 0.75], [1e10, -1e10, 123.456]]
    seeds = np.array([0, 12345, -2**63, 2**63 - 1], np.int64)
    names = OW_NOISES if dim == "overworld" else NE_NOISES
    for k, name in enumerate(names):
        got = eng.noise_values(k, seeds, xyz)
        for i, s in enumerate(seeds):
            want = np.array([orc.noise(name, int(s), *p) for p in xyz])
            assert np.array_equal(got[i].view(np.uint64), want.view(np.uint64)), (name, int(s))
