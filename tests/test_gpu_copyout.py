"""Results copied back to the caller's memory through the staged copy (blocks of pinned memory, host
threads moving each block out while the next one travels): sizes that are not multiples of a block or of
a page, a destination that starts off a page boundary, sizes around the 1 MiB plain-copy threshold, and
back-to-back calls of different sizes, and a pinned destination (one direct copy).  Each output is
checked against the per-point paths, whose results are small copies, and bytes around an unaligned
destination must stay untouched."""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
mcgpu = pytest.importorskip("mcgpu")
mcbiome = pytest.importorskip("mcbiome")


@pytest.fixture(scope="module")
def world():
    if mcgpu.lib().mcgpu_device_count() < 1:
        pytest.fail("no HIP device visible to libmcbiome_hip.so")
    o = oracle.Oracle("26.3", "overworld", "overworld")
    ref = oracle.McgWorldRef.vanilla(o, 4242)
    W = C.cast(C.c_void_p(ref.addr), C.POINTER(mcgpu.McgBiomeWorld)).contents
    return ref, mcgpu.BiomeWorld(W)


def _grid_into(w, buf, off, x0, z0, nx, nz, step, y):
    bad = C.c_int(0)
    p = C.cast(C.c_void_p(buf.ctypes.data + off), C.POINTER(C.c_uint8))
    rc = mcgpu.lib().mcgpu_biome_grid(w.h, x0, z0, nx, nz, step, y, p, C.byref(bad))
    assert rc == 0, mcgpu.last_error()
    return bad.value


def _sample_check(w, g, x0, z0, nx, nz, step, y, seed):
    rng = np.random.default_rng(seed)
    i = rng.integers(0, nx, 40_000)
    j = rng.integers(0, nz, 40_000)
    # the corners and the last row/column (the end of the last block) always
    i = np.concatenate([i, [0, nx - 1, 0, nx - 1], np.arange(nx)])
    j = np.concatenate([j, [0, 0, nz - 1, nz - 1], np.full(nx, nz - 1)])
    xyz = np.stack([x0 + i * step, np.full(i.size, y), z0 + j * step], 1).astype(np.int32)
    p, _ = w.points(xyz)
    assert np.array_equal(g[j, i], p), int((g[j, i] != p).sum())


@pytest.mark.parametrize("nx,nz,off", [(3001, 1799, 3), (1024, 1023, 4095), (1024, 1024, 0), (1031, 1019, 1),
                                        (4099, 2053, 17)])
def test_plugin_grid_unaligned_destination(world, nx, nz, off):
    ref, w = world
    x0, z0, step, y = -5003, 7001, 1, 64
    n = nx * nz
    buf = np.full(n + off + 4096, 0xA5, np.uint8)
    print("grid", nx, nz, off, flush=True)
    bad = _grid_into(w, buf, off, x0, z0, nx, nz, step, y)
    assert (buf[:off] == 0xA5).all() and (buf[off + n:] == 0xA5).all()
    g = buf[off:off + n].reshape(nz, nx)
    a, bad_a = w.grid(x0, z0, nx, nz, step, y)
    assert bad == bad_a and np.array_equal(g, a)
    _sample_check(w, g, x0, z0, nx, nz, step, y, nx)


def test_back_to_back_sizes(world):
    """Large, small, large again and a size just over the threshold: each call's bytes land whole."""
    ref, w = world
    for k, (nx, nz) in enumerate([(2900, 2111), (300, 200), (3500, 3001), (1025, 1024), (2048, 2049)]):
        x0, z0 = -70_000 + 977 * k, 30_000 - 1201 * k
        print("call", k, nx, nz, flush=True)
        g, _ = w.grid(x0, z0, nx, nz, 1, 70)
        _sample_check(w, g, x0, z0, nx, nz, 1, 70, k)


def test_region_and_chunks_against_points():
    """mcb_region (a one-off plan) and mcb_chunk_biomes return through the staged copy too."""
    eng = mcbiome.Engine("26.3", "overworld")
    seed, x0, z0, nx, nz, ys = 77, -1000, 500, 1201, 1003, [0, 16, 31]
    print("region", nx, nz, len(ys), flush=True)
    r = eng.region(seed, x0, z0, nx, nz, ys)
    rng = np.random.default_rng(5)
    k = rng.integers(0, len(ys), 30_000)
    i = rng.integers(0, nx, 30_000)
    j = rng.integers(0, nz, 30_000)
    i[:2], j[:2], k[:2] = nx - 1, nz - 1, len(ys) - 1
    pts = np.stack([x0 + i, np.asarray(ys)[k], z0 + j], 1).astype(np.int32)
    p = eng.points([seed], pts)[0]
    assert np.array_equal(r[k, j, i], p), int((r[k, j, i] != p).sum())
    xs, zs = np.meshgrid(np.arange(70), np.arange(70))
    chunks = np.stack([300 + xs.ravel(), -900 + zs.ravel()], 1).astype(np.int32)
    print("chunks", len(chunks), flush=True)
    c = eng.chunk_biomes(seed, chunks)
    sel = rng.integers(0, len(chunks), 40)
    sub = eng.chunk_biomes(seed, chunks[sel])
    assert np.array_equal(c[sel], sub)


def test_pinned_destination(world):
    """A destination in pinned host memory (mcgpu_host_alloc, as libmcgen's terrain service uses) takes one
    direct copy; the bytes equal those of a pageable destination, and the bytes around it stay untouched."""
    ref, w = world
    nx, nz, off = 2500, 1700, 64
    n = nx * nz
    L = mcgpu.lib()
    p = L.mcgpu_host_alloc(n + 2 * off)
    assert p
    try:
        buf = np.ctypeslib.as_array(C.cast(C.c_void_p(p), C.POINTER(C.c_uint8)), shape=(n + 2 * off,))
        buf[:] = 0x5A
        print("pinned grid", nx, nz, flush=True)
        bad = _grid_into(w, buf, off, 900, -1200, nx, nz, 1, 64)
        a, bad_a = w.grid(900, -1200, nx, nz, 1, 64)
        assert bad == bad_a and np.array_equal(buf[off:off + n].reshape(nz, nx), a)
        assert (buf[:off] == 0x5A).all() and (buf[off + n:] == 0x5A).all()
    finally:
        L.mcgpu_host_free(p)
