"""The libmcgen-compatible biome API on the device: mcgen_biome_grid / mcgen_biome_at /
mcgen_gpu_biome_points against the restatement of libmcgen's mcgen_biome_at (uncapped tie chains), and
mcgen_generate_region(MC_STAGE_BIOMES) against the restatement of world_chunk_biomes; the self-test."""
import ctypes as C

import numpy as np
import pytest

import mcgprog
import oracle as O

pytestmark = pytest.mark.gpu
mcgen = pytest.importorskip("mcgen")
DIMS = {"overworld": "minecraft:overworld", "nether": "minecraft:the_nether", "end": "minecraft:the_end"}


def _to_engine_ids(g):
    """libmcgen id (the version's registry) -> the engine's id of the same biome name."""
    return np.array([O.biome_id(n.split(":", 1)[1]) for n in g.biome_names()], np.int64)


@pytest.fixture(scope="module", autouse=True)
def _device():
    if mcgen.lib().mcgen_gpu_device_count() < 1:
        pytest.fail("no HIP device visible to libmcbiome_hip.so")


@pytest.mark.parametrize("version", ["26.1", "26.3"])
@pytest.mark.parametrize("dim", ["overworld", "nether", "end"])
def test_biomes_match_libmcgen_semantics(version, dim):
    seed = 4242
    g = mcgen.Gen(version)
    w = g.world(DIMS[dim], "normal", seed)
    V = mcgprog.vanilla(version, dim, seed)
    ids = _to_engine_ids(g)
    print("mcgen", version, dim, flush=True)
    for (x0, z0, n, step, y) in [(-700, 333, 96, 1, 64), (12000, -90000, 64, 5, 30), (0, 0, 40, 16, 500)]:
        got = w.biome_grid(x0, z0, n + 3, n, step, y)
        want, bad = V.ref.grid(x0, z0, n + 3, n, step, y, cap=-1)
        assert bad == 0 and np.array_equal(ids[got], want), (x0, z0, step)
    rng = np.random.default_rng(5)
    xyz = np.stack([rng.integers(-10**6, 10**6, 300), rng.integers(-100, 400, 300), rng.integers(-10**6, 10**6, 300)], 1)
    want, _ = V.ref.points(xyz, cap=-1)
    assert [int(ids[w.biome_at(*map(int, p))]) for p in xyz[:20]] == [int(b) for b in want[:20]]
    out = np.empty(len(xyz), np.uint8)
    ncpu = C.c_int(-1)
    xyz32 = np.ascontiguousarray(xyz, np.int32)
    assert mcgen.lib().mcgen_gpu_biome_points(w.h, len(xyz), xyz32.ctypes.data_as(C.POINTER(C.c_int)),
                                               out.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(ncpu)) == 0
    assert ncpu.value == 0 and np.array_equal(ids[out], want)
    assert (mcgen.lib().mcgen_world_min_y(w.h), mcgen.lib().mcgen_world_height(w.h)) == \
        ((-64, 384) if dim == "overworld" else (0, 256))
    w.close()


@pytest.mark.parametrize("version,dim", [("26.3", "overworld"), ("26.2", "nether"), ("26.1", "end")])
def test_region_biomes_match_world_chunk_biomes(version, dim):
    seed = -99
    g = mcgen.Gen(version)
    w = g.world(DIMS[dim], "normal", seed)
    got = w.region_biomes(-3, 17, 5, 4)
    V = mcgprog.vanilla(version, dim, seed)
    chunks = np.array([[-3 + ix, 17 + iz] for iz in range(4) for ix in range(5)], np.int32)
    want = V.oracle.chunk_biomes(seed, chunks)
    assert np.array_equal(_to_engine_ids(g)[got.reshape(want.shape)], want)
    r = C.c_void_p()
    rc = mcgen.lib().mcgen_generate_region(w.h, 0, 0, 1, 1, mcgen.MC_STAGE_BIOMES | mcgen.MC_STAGE_TERRAIN, 0,
                                           mcgen.PROGRESS_FN(0), None, C.byref(r), None, 0)
    assert rc == mcgen.MCGEN_E_UNSUPPORTED


def test_selftest_and_status():
    g = mcgen.Gen("26.3")
    rep = C.create_string_buffer(1024)
    assert mcgen.lib().mcgen_gpu_selftest(g.h, 400, rep, 1024) == 0, rep.value
    print(rep.value.decode(), flush=True)
    st = C.create_string_buffer(512)
    assert mcgen.lib().mcgen_gpu_status(st, 512) == 1 and b"GPU ready" in st.value
