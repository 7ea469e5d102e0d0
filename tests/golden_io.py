"""Readers for the golden fixtures under tests/golden/.

vectors/<V>/*      : the reference's own Java-oracle vectors (tests/vectors/<V>/ in the reference),
                     copied verbatim as data; misc_biome.jsonl keeps only its hashed/blockbiome records.
unit/ref26{2,3}.txt: the reference's NormalNoise unit KATs (tests/unit/ in the reference).
engine/*.npz       : outputs of the reference CPU engine built from its sources (oracle/_ref),
                     written by tests/golden/make_engine_fixtures.py.
"""
import json
import os
import struct
from collections import defaultdict

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
VERSIONS = ("26.1", "26.2", "26.3")


def _rows(path):
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            yield line.rstrip("\n").split("\t")


def bits_to_double(b):
    return struct.unpack("<d", struct.pack("<q", int(b)))[0]


def f32_bits(v):
    return struct.unpack("<I", struct.pack("<f", v))[0]


def f64_bits(v):
    return struct.unpack("<q", struct.pack("<d", v))[0]


def climate_rows(version, dim):
    """[(seed, qx, qy, qz, [t,h,c,e,d,w], biome_name)] for overworld / the_nether;
    End rows: (seed, qx, qy, qz, height_value, fbits, biome_name)."""
    out = []
    for r in _rows(os.path.join(GOLDEN, "vectors", version, "climate_%s.tsv" % dim)):
        if dim == "the_end":
            out.append((int(r[0]), int(r[1]), int(r[2]), int(r[3]), float(r[4]), int(r[5]), r[6].replace("minecraft:", "")))
        else:
            out.append((int(r[0]), int(r[1]), int(r[2]), int(r[3]), [int(x) for x in r[4:10]], r[10].replace("minecraft:", "")))
    return out


def by_seed(rows):
    g = defaultdict(list)
    for r in rows:
        g[r[0]].append(r)
    return g


def noise_rows(version):
    """[(dim, seed, noise_name, x, y, z, value, dbits, fbits)]"""
    out = []
    for r in _rows(os.path.join(GOLDEN, "vectors", version, "noise.tsv")):
        out.append((r[0], int(r[1]), r[2].replace("minecraft:", ""), float(r[3]), float(r[4]), float(r[5]),
                    float(r[6]), int(r[7]), int(r[8])))
    return out


def df_rows(version):
    out = []
    for r in _rows(os.path.join(GOLDEN, "vectors", version, "df.tsv")):
        out.append((r[0], int(r[1]), r[2].replace("minecraft:", ""), int(r[3]), int(r[4]), int(r[5]),
                    float(r[6]), int(r[7]), int(r[8])))
    return out


def misc_records(version, cmd):
    out = []
    with open(os.path.join(GOLDEN, "vectors", version, "misc_biome.jsonl")) as f:
        for line in f:
            d = json.loads(line)
            if d.get("cmd") == cmd:
                out.append(d)
    return out


def grid_hashes(version):
    """[(dim, seed, mode, qx0, qz0, n, qy, hash_int)]"""
    out = []
    for r in _rows(os.path.join(GOLDEN, "vectors", version, "biome_grid_hashes.tsv")):
        out.append((r[0], int(r[1]), r[2], int(r[3]), int(r[4]), int(r[5]), int(r[6]), int(r[7], 16)))
    return out


def unit_kat(version):
    """[(seed, noise, x, y, z, bits)] from ref263.txt (float bits) / ref262.txt (double bits)."""
    name = "ref263.txt" if version == "26.3" else "ref262.txt"
    out = []
    with open(os.path.join(GOLDEN, "unit", name)) as f:
        for line in f:
            s, nm, xb, yb, zb, vb = line.split()
            out.append((int(s), nm, bits_to_double(xb), bits_to_double(yb), bits_to_double(zb), int(vb, 16)))
    return out


# NormalNoise parameters of the unit KATs (tests/unit/test_noise_new.c:6-13, test_noise_old.c:7-14)
UNIT_NOISES = {
    "temperature": (-10, [1.5, 0, 1, 0, 0, 0], 1.2453007926713473),
    "vegetation": (-8, [1, 1, 0, 0, 0, 0], 0.9494731054427978),
    "continentalness": (-9, [1, 1, 2, 2, 2, 1, 1, 1, 1], 0.8880832896205223),
    "erosion": (-9, [1, 1, 0, 1, 1], 1.063180125160734),
    "ridge": (-7, [1, 2, 1, 0, 0, 0], 0.9147152149950137),
    "offset": (-3, [1, 1, 1, 0], 0.9381732587751005),
}
