"""Pin the oracle against the reference's known-answer tests (bit patterns from the Java oracle).

Covers: Xoroshiro seeding/positional forks, MD5 name hashes, octave init, ImprovedNoise /
GradientNoise, NormalNoise (both stack orders), legacy Nether init, the climate sampler with
spline and depth, SHA-256 hashed seed and BiomeManager zoom.
"""
import numpy as np
import pytest

import golden_io as G
import oracle as O


@pytest.mark.parametrize("version", ["26.2", "26.3"])
def test_unit_normalnoise_bits(version):
    """tests/unit/ref26{2,3}.txt: 9 seeds x 6 noises x 40 points, exact bits (2160 each)."""
    bad = 0
    rows = G.unit_kat(version)
    assert len(rows) == 2160
    for seed, nm, x, y, z, bits in rows:
        fo, amps, base = G.UNIT_NOISES[nm]
        v = O.noise_value_params(version, nm, fo, amps, base, seed, x, y, z)
        got = G.f32_bits(v) if version == "26.3" else G.f64_bits(v) & (2**64 - 1)
        bad += got != bits
    assert bad == 0


CLIMATE_NOISES = {"temperature", "vegetation", "continentalness", "erosion", "ridge", "offset"}


@pytest.mark.parametrize("version", G.VERSIONS)
def test_noise_tsv_climate_rows(version):
    """vectors/<V>/noise.tsv: the overworld climate noises and the legacy Nether biome noises."""
    worlds = {"overworld": O.Oracle(version, "overworld"), "the_nether": O.Oracle(version, "the_nether")}
    n = bad = 0
    for dim, seed, name, x, y, z, value, dbits, fbits in G.noise_rows(version):
        if dim == "overworld" and name in CLIMATE_NOISES:
            w = worlds["overworld"]
        elif dim == "the_nether" and name.startswith("nether/"):
            w = worlds["the_nether"]
        else:
            continue
        v = w.noise(name, seed, x, y, z)
        n += 1
        if version == "26.3":
            bad += G.f32_bits(v) != (fbits & 0xFFFFFFFF)
        else:
            bad += G.f64_bits(v) != dbits
    assert n == 400 and bad == 0


DF_INDEX = {"overworld/continents": 0, "overworld/erosion": 1, "overworld/ridges": 2, "overworld/offset": 3,
            "overworld/depth": 4}


@pytest.mark.parametrize("version", G.VERSIONS)
def test_df_tsv(version):
    """vectors/<V>/df.tsv: continents/erosion/ridges/offset/depth density functions at block points."""
    w = O.Oracle(version, "overworld")
    n = bad = 0
    for dim, seed, name, x, y, z, value, dbits, fbits in G.df_rows(version):
        v = w.climate_df(seed, x, y, z)[DF_INDEX[name]]
        n += 1
        if version == "26.3":
            bad += G.f32_bits(v) != (fbits & 0xFFFFFFFF)
        else:
            bad += G.f64_bits(v) != dbits
    assert n == 200 and bad == 0


@pytest.mark.parametrize("version", G.VERSIONS)
def test_hashed_seed(version):
    recs = G.misc_records(version, "hashed")
    assert len(recs) == 20
    for r in recs:
        assert O.hashed_seed(r["seed"]) == r["hashed_seed"]


def test_sha256_textbook():
    """FIPS 180-4 via hashlib: the first 8 digest bytes, read little-endian, are the hashed seed."""
    import hashlib
    rng = np.random.default_rng(5)
    for s in [0, 1, -1, 2**63 - 1, -2**63] + [int(v) for v in rng.integers(-2**63, 2**63 - 1, 50, dtype=np.int64)]:
        d = hashlib.sha256(int(s).to_bytes(8, "little", signed=True)).digest()
        assert O.hashed_seed(s) == int.from_bytes(d[:8], "little", signed=True)


@pytest.mark.parametrize("version", G.VERSIONS)
def test_blockbiome_zoom(version):
    """misc_biome.jsonl blockbiome records: BiomeManager zoom + climate + tree at block positions."""
    recs = G.misc_records(version, "blockbiome")
    assert len(recs) == 30
    names = O.biome_names()
    worlds = {}
    for r in recs:
        key = (r["dim"], r.get("preset", "normal"))
        if key not in worlds:
            worlds[key] = O.Oracle(version, r["dim"], r.get("preset", "normal"))
        w = worlds[key]
        assert O.hashed_seed(r["seed"]) == r["hashed_seed"]
        b = w.block_grid(r["seed"], r["x"], r["z"], 1, 1, 1, r["y"])[0, 0]
        assert names[b] == r["biome"].replace("minecraft:", ""), r


def test_zoom_brute_force():
    """The zoom picks the lattice corner minimising the fiddled distance; restated here in
    Python integers/floats (an independent formula) on random blocks."""
    rng = np.random.default_rng(7)
    M = 2**64

    def step(r, c):
        v = r % M
        v = (v * ((v * 6364136223846793005 + 1442695040888963407) % M) + c) % M
        return v - M if v >= 2**63 else v

    def fid(r):
        return (((r >> 24) % 1024) / 1024.0 - 0.5) * 0.9

    for _ in range(300):
        hs = int(rng.integers(-2**63, 2**63 - 1, dtype=np.int64))
        x, y, z = (int(v) for v in rng.integers(-2**22, 2**22, 3))
        ax, ay, az = x - 2, y - 2, z - 2
        base = (ax >> 2, ay >> 2, az >> 2)
        frac = ((ax & 3) / 4.0, (ay & 3) / 4.0, (az & 3) / 4.0)
        best, bd = None, float("inf")
        for c in range(8):
            u = ((c >> 2) & 1, (c >> 1) & 1, c & 1)
            cell = tuple(b + k for b, k in zip(base, u))
            r = hs
            for v in cell + cell:
                r = step(r, v)
            o = [fid(r)]
            r = step(r, hs); o.append(fid(r))
            r = step(r, hs); o.append(fid(r))
            dd = [f - k + off for f, k, off in zip(frac, u, o)]
            d = dd[2] * dd[2] + dd[1] * dd[1] + dd[0] * dd[0]
            if bd > d:
                best, bd = cell, d
        assert O.zoom(hs, x, y, z) == best
