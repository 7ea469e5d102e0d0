"""The plugin's vanilla-router path (the engine's fused column code running a recognised vanilla Overworld
program) against the density-function interpreter on the same world, and against the restatement of
libmcgen's biome lookup (oracle/mcgplugin.c).

* every version and preset: grids through the cell lattice (steps 1 and 4), the per-point grid path
  (step 16) and points, at ordinary coordinates and past +-2^26 blocks (PerlinNoise.wrap active), with y
  beyond the world's height; outputs and unresolved counts equal between the two paths and the oracle;
* a world whose spline coordinate list is permuted takes the router with the coordinates mapped;
* a perturbed world (blend alpha 0.5) is interpreted, and still equals the oracle.
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


def _interpreted(world):
    os.environ["MCB_PLUGIN_INTERPRET"] = "1"
    try:
        return mcgpu.BiomeWorld(world)
    finally:
        del os.environ["MCB_PLUGIN_INTERPRET"]


def _world(version, preset, seed):
    o = oracle.Oracle(version, "overworld", preset)
    ref = oracle.McgWorldRef.vanilla(o, seed)
    W = C.cast(C.c_void_p(ref.addr), C.POINTER(mcgpu.McgBiomeWorld)).contents
    return ref, W


CASES = [(v, p) for v in ("26.1", "26.2", "26.3") for p in ("overworld", "large_biomes", "amplified")]
# (step, x0, z0, n, y): y -64 puts the lattice's lower layer on the world's lowest quart, where a chunk's
# first fill cell settles its ties without a chain; y 330 clamps both layers to the top
GRIDS = [(1, -1203, 2051, 96, 70), (4, 300_001, -77_777, 64, -40), (16, -9_000_000, 4_000_000, 48, 200),
         (1, 2**28 + 5, -(2**28) - 11, 40, 64), (4, -(2**30) + 1000, 2**30 - 5000, 40, 330), (1, 5003, -9001, 120, -64),
         (2, -77, 31, 200, -61)]


@pytest.mark.parametrize("version,preset", CASES)
def test_router_equals_interpreter_and_oracle(version, preset):
    seed = zlib.crc32(("%s/%s" % (version, preset)).encode()) - 2**31
    ref, W = _world(version, preset, seed)
    A = mcgpu.BiomeWorld(W)
    B = _interpreted(W)
    assert A.router and not B.router
    for step, x0, z0, n, y in GRIDS:
        print("grid", version, preset, step, x0, z0, n, y, flush=True)
        ga, ba = A.grid(x0, z0, n + 3, n, step, y)
        gb, bb = B.grid(x0, z0, n + 3, n, step, y)
        assert ba == bb and np.array_equal(ga, gb), int((ga != gb).sum())
        gw, bw = ref.grid(x0, z0, n + 3, n, step, y)
        assert ba == bw and np.array_equal(ga, gw), int((ga != gw).sum())
    rng = np.random.default_rng(zlib.crc32(version.encode() + preset.encode()))
    xyz = np.stack([rng.integers(-2**30, 2**30, 2000), rng.integers(-120, 400, 2000),
                    rng.integers(-2**30, 2**30, 2000)], 1).astype(np.int32)
    xyz = np.concatenate([xyz, np.stack([rng.integers(-5000, 5000, 2000), rng.integers(-64, 320, 2000),
                                         rng.integers(-5000, 5000, 2000)], 1).astype(np.int32)])
    print("points", version, preset, len(xyz), flush=True)
    pa, qa = A.points(xyz)
    pb, qb = B.points(xyz)
    pw, qw = ref.points(xyz)
    assert qa == qb == qw
    assert np.array_equal(pa, pb) and np.array_equal(pa, pw), (int((pa != pb).sum()), int((pa != pw).sum()))


def test_router_lattice_large_grid_equals_interpreter():
    """The lattice launches over a 4096 x 1024 grid (the plugin benchmark's shape: ~950 tied cells go to the
    state machine) on both paths."""
    ref, W = _world("26.3", "overworld", 12345)
    A = mcgpu.BiomeWorld(W)
    B = _interpreted(W)
    print("large grid", flush=True)
    ga, ba = A.grid(-2048, -512, 4096, 1024, 1, 64)
    gb, bb = B.grid(-2048, -512, 4096, 1024, 1, 64)
    assert ba == bb and np.array_equal(ga, gb), int((ga != gb).sum())


@pytest.mark.parametrize("version", ["26.3", "26.1"])
def test_router_points_path_large_grid_equals_interpreter(version):
    """A step-4 grid (one zoomed cell per block: the per-point launches, the shape of libmcgen's biome-map
    benchmark) of 1024 x 768 on both paths, and points with ties near the world's bottom."""
    ref, W = _world(version, "overworld", 12345)
    A = mcgpu.BiomeWorld(W)
    B = _interpreted(W)
    print("step-4 grid", version, flush=True)
    ga, ba = A.grid(-2048, -1536, 1024, 768, 4, 64)
    gb, bb = B.grid(-2048, -1536, 1024, 768, 4, 64)
    assert ba == bb and np.array_equal(ga, gb), int((ga != gb).sum())
    rng = np.random.default_rng(31)
    xyz = np.stack([rng.integers(-300_000, 300_000, 200_000), rng.integers(-70, -50, 200_000),
                    rng.integers(-300_000, 300_000, 200_000)], 1).astype(np.int32)
    pa, qa = A.points(xyz)
    pb, qb = B.points(xyz)
    assert qa == qb and np.array_equal(pa, pb), int((pa != pb).sum())


def test_router_with_permuted_spline_coordinates():
    ref, W = _world("26.3", "overworld", 99)
    P = W.prog.contents
    s = [P.nodes[i] for i in range(P.nnodes) if P.nodes[i].k == mcgpu.K["SPLINE"]][0]
    perm = [3, 1, 0, 2]
    old = [P.arr[s.arr0 + j] for j in range(4)]
    for j in range(4):
        P.arr[s.arr0 + perm[j]] = old[j]
    for i in range(P.nsp):
        if not P.sp[i].is_const:
            P.sp[i].coord = perm[P.sp[i].coord]
    A = mcgpu.BiomeWorld(W)
    assert A.router
    ga, ba = A.grid(-700, 900, 120, 100, 2, 64)
    gw, bw = ref.grid(-700, 900, 120, 100, 2, 64)
    assert ba == bw and np.array_equal(ga, gw)


def test_perturbed_world_is_interpreted():
    ref, W = _world("26.3", "overworld", 5)
    P = W.prog.contents
    al = [i for i in range(P.nnodes) if P.nodes[i].k == mcgpu.K["CTX_ALPHA"]][0]
    P.nodes[al].f0 = 0.5
    A = mcgpu.BiomeWorld(W)
    assert not A.router
    ga, ba = A.grid(1000, -3000, 90, 80, 1, 64)
    gw, bw = ref.grid(1000, -3000, 90, 80, 1, 64)
    assert ba == bw and np.array_equal(ga, gw)


@pytest.mark.parametrize("version", ["26.3", "26.1"])
def test_router_chunks_equal_interpreter(version):
    """Whole chunks (mcb_chunk_biomes: every quart layer, ties along the fill order with no cap) through
    the router launches and through the interpreter."""
    import mcbiome
    eng = mcbiome.Engine(version, "overworld")
    xs, zs = np.meshgrid(np.arange(24), np.arange(20))
    chunks = np.stack([-4000 + xs.ravel(), 9000 + zs.ravel()], 1).astype(np.int32)
    print("chunks", version, len(chunks), flush=True)
    a = eng.chunk_biomes(77, chunks)
    os.environ["MCB_PLUGIN_INTERPRET"] = "1"
    try:
        b = eng.chunk_biomes(77, chunks)
    finally:
        del os.environ["MCB_PLUGIN_INTERPRET"]
    assert np.array_equal(a, b), int((a != b).sum())
