"""Readers for the golden fixtures under tests/golden/.

vectors/<V>/*      : the reference's own Java-oracle vectors (tests/vectors/<V>/ in the reference),
                     copied verbatim as data; misc_biome.jsonl keeps only its hashed/blockbiome records.
unit/ref26{2,3}.txt: the reference's NormalNoise unit KATs (tests/unit/ in the reference).
engine/*.npz       : outputs of the reference CPU engine built from its sources (oracle/_ref) at the
                     benchmark workloads' own inputs, written by tests/golden/make_engine_fixtures.py
                     (plain_26.3: the one-seed 256x256 grid; batched_<V>: per-seed hashes of the
                     4096-seed batch).
"""
import json
import os
import struct
from collections import defaultdict

import numpy as np

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


# ---- splitmix64 workloads and the per-seed output hash of the engine fixtures ----
MASK64 = (1 << 64) - 1


def splitmix64(k, n):
    """First n outputs (unsigned) of the splitmix64 stream with state0 = k: state += 0x9E3779B97F4A7C15,
    then the finalizer of engine/mc_rng.h:64-68 (the 30/27/31 shifts)."""
    st, out = k & MASK64, []
    for _ in range(n):
        st = (st + 0x9E3779B97F4A7C15) & MASK64
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        out.append(z ^ (z >> 31))
    return out


def signed64(v):
    return v - (1 << 64) if v >= (1 << 63) else v


def uniform(d, r):
    """Uniform integer in [-r, r] from one unsigned draw."""
    return d % (2 * r + 1) - r


def fnv1a64_rows(a):
    """FNV-1a-64 of each row of a 2-D uint8 array (offset basis 0xcbf29ce484222325, prime 0x100000001b3)."""
    a = np.ascontiguousarray(a, dtype=np.uint8)
    h = np.full(a.shape[0], 0xcbf29ce484222325, np.uint64)
    prime = np.uint64(0x100000001b3)
    with np.errstate(over="ignore"):
        for k in range(a.shape[1]):
            h ^= a[:, k].astype(np.uint64)
            h *= prime
    return h


def engine_fixture(name):
    """tests/golden/engine/<name>.npz as a dict of arrays."""
    with np.load(os.path.join(GOLDEN, "engine", name + ".npz")) as z:
        return {k: z[k] for k in z.files}
