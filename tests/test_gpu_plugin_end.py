"""The plugin's End table (k_end_table: the islands heights of a grid's lattice of chunks, each tile's
islands listed once) against the interpreter computing every chunk's 25 x 25 window
(MCB_PLUGIN_INTERPRET=1), and against the restatement of libmcgen's End biome source (oracle/mcgplugin.c).

* 26.1 / 26.2: the islands node with its centre term, whose square wraps to a negative int and gives NaN
  far from the origin (the End biome is then the last of the five);
* 26.3: the maximum of a distance term and the islands node;
* grids over the central island and its ring, across chunk 0 (where a section's window centre s / 2 jumps
  under C division), partial tiles, and far out; steps 1, 2 and 4 (the lattice path).
"""
import ctypes as C
import os
import zlib

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
mcgpu = pytest.importorskip("mcgpu")


@pytest.fixture(scope="module", autouse=True)
def _device():
    if mcgpu.lib().mcgpu_device_count() < 1:
        pytest.fail("no HIP device visible to libmcbiome_hip.so")


def _worlds(version, seed):
    o = oracle.Oracle(version, "the_end")
    ref = oracle.McgWorldRef.vanilla(o, seed)
    W = C.cast(C.c_void_p(ref.addr), C.POINTER(mcgpu.McgBiomeWorld)).contents
    A = mcgpu.BiomeWorld(W)
    os.environ["MCB_PLUGIN_INTERPRET"] = "1"
    try:
        B = mcgpu.BiomeWorld(W)
    finally:
        del os.environ["MCB_PLUGIN_INTERPRET"]
    return ref, A, B


# (step, x0, z0, nx, nz, y, check_with_oracle)
GRIDS = [(1, -700, -650, 1500, 1400, 64, False),           # the central island, its ring, chunks -1 / 0
         (4, -1500, -1300, 130, 110, 0, True),
         (2, 5, 5, 64, 64, 64, True),                       # inside the central island only
         (4, -(2**20) - 77, 3 * 2**19 + 5, 700, 500, 40, False),
         (1, 3000, -2000, 97, 83, 64, True),                 # partial tiles
         (4, 2**26 - 3000, -(2**26) + 2000, 600, 500, 64, False),   # 26.1/26.2: the centre term is NaN here
         (4, -(2**27) + 64, 2**27 - 4000, 90, 70, 64, True)]


@pytest.mark.parametrize("version", ["26.1", "26.2", "26.3"])
def test_end_table_equals_interpreter_and_oracle(version):
    seed = zlib.crc32(("end/%s" % version).encode()) - 2**31
    ref, A, B = _worlds(version, seed)
    for step, x0, z0, nx, nz, y, with_oracle in GRIDS:
        print("end grid", version, step, x0, z0, nx, nz, flush=True)
        ga, ba = A.grid(x0, z0, nx, nz, step, y)
        gb, bb = B.grid(x0, z0, nx, nz, step, y)
        assert ba == bb == 0 and np.array_equal(ga, gb), int((ga != gb).sum())
        if with_oracle:
            gw, bw = ref.grid(x0, z0, nx, nz, step, y)
            assert bw == 0 and np.array_equal(ga, gw), int((ga != gw).sum())


def test_end_table_large_grid():
    """The libmcgen biome-map benchmark's shape in the End: 2048 x 2048 samples at step 4 (512 x 512 chunks,
    1,024 tiles)."""
    ref, A, B = _worlds("26.3", 12345)
    print("end benchmark grid", flush=True)
    ga, _ = A.grid(-4096, -4096, 2048, 2048, 4, 64)
    gb, _ = B.grid(-4096, -4096, 2048, 2048, 4, 64)
    assert np.array_equal(ga, gb), int((ga != gb).sum())


@pytest.mark.parametrize("version", ["26.1", "26.3"])
def test_end_table_points_and_sparse_grids(version):
    """Points and grids sampled more sparsely than one block per chunk (the per-point path) read the table
    over the chunks their zoomed quarts can reach when it is small enough; scattered points compute their
    windows."""
    ref, A, B = _worlds(version, 4242)
    rng = np.random.default_rng(7)
    clustered = np.stack([rng.integers(-3000, 3000, 5000), rng.integers(0, 256, 5000),
                          rng.integers(-2500, 2500, 5000)], 1).astype(np.int32)
    scattered = np.stack([rng.integers(-2**30, 2**30, 300), rng.integers(0, 256, 300),
                          rng.integers(-2**30, 2**30, 300)], 1).astype(np.int32)
    for name, xyz in (("clustered", clustered), ("scattered", scattered)):
        print("end points", version, name, flush=True)
        pa, qa = A.points(xyz)
        pb, qb = B.points(xyz)
        pw, qw = ref.points(xyz)
        assert qa == qb == qw == 0
        assert np.array_equal(pa, pb) and np.array_equal(pa, pw), (int((pa != pb).sum()), int((pa != pw).sum()))
    for step, x0, z0, nx, nz in [(16, -3001, 1999, 300, 200), (64, 2**24 + 3, -(2**24) - 9, 150, 100)]:
        print("end sparse grid", version, step, flush=True)
        ga, _ = A.grid(x0, z0, nx, nz, step, 64)
        gb, _ = B.grid(x0, z0, nx, nz, step, 64)
        gw, _ = ref.grid(x0, z0, nx, nz, step, 64)
        assert np.array_equal(ga, gb) and np.array_equal(ga, gw), (int((ga != gb).sum()), int((ga != gw).sum()))


@pytest.mark.parametrize("version", ["26.2", "26.3"])
def test_end_table_chunk_biomes(version):
    """Whole End chunks (mcb_chunk_biomes): every quart cell of a chunk evaluates at the chunk's section,
    read from the table over the listed chunks' rectangle."""
    import mcbiome
    eng = mcbiome.Engine(version, "the_end")
    xs, zs = np.meshgrid(np.arange(40), np.arange(30))
    near = np.stack([-70 + xs.ravel(), -10 + zs.ravel()], 1).astype(np.int32)       # the ring of the centre
    far = np.stack([3 * 2**20 + xs.ravel(), -(2**21) + zs.ravel()], 1).astype(np.int32)
    for chunks in (near, far, np.concatenate([near[:50], far[:50]])):
        print("end chunks", version, len(chunks), flush=True)
        a = eng.chunk_biomes(99, chunks)
        os.environ["MCB_PLUGIN_INTERPRET"] = "1"
        try:
            b = eng.chunk_biomes(99, chunks)
        finally:
            del os.environ["MCB_PLUGIN_INTERPRET"]
        assert np.array_equal(a, b), int((a != b).sum())
