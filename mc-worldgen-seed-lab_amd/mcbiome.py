"""Python binding of libmcbiome_hip.so (the MI355X biome engine), over its C ABI (include/mcbiome.h).

Host-side mirror of the reference's CPU-engine interface (engine/cuda/mcg_host.h): ``Host`` loads a
(version, dimension, preset) world like ``mcg_host_load`` and ``Host.ref_points`` has the contract of
``mcg_ref_points``.  ``Engine`` exposes the native batched API (device buffers, plans, hipGraph).

There is no CPU fallback: importing this module fails loudly when the HIP library is not built, and
every compute call raises ``EngineError`` when the device is unusable.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmcbiome_hip.so")
DATA_DIR = os.path.join(HERE, "data")

VERSIONS = {"26.1": 0, "26.2": 1, "26.3": 2, "26.4-snapshot-2": 3}
DIMS = {"overworld": 0, "the_nether": 1, "nether": 1, "the_end": 2, "end": 2}


class PluginTerrainWorld:
    """Owner of a McgTerrainWorld made by mcb_plugin_terrain_world; ``.world`` is the mcgpu.McgTerrainWorld."""

    def __init__(self, engine, seed):
        import mcgpu
        p = C.c_void_p()
        _check(lib().mcb_plugin_terrain_world(engine.h, int(seed), C.byref(p)))
        self.addr = p.value
        self.world = C.cast(p, C.POINTER(mcgpu.McgTerrainWorld)).contents
        self._engine = engine

    def close(self):
        if getattr(self, "addr", None):
            lib().mcb_plugin_terrain_world_free(self.addr)
            self.addr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PluginWorld:
    """Owner of a McgBiomeWorld made by mcb_plugin_world."""

    def __init__(self, engine, seed):
        p = C.c_void_p()
        _check(lib().mcb_plugin_world(engine.h, int(seed), C.byref(p)))
        self.addr = p.value
        self._engine = engine

    def close(self):
        if getattr(self, "addr", None):
            lib().mcb_plugin_world_free(self.addr)
            self.addr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EngineError(RuntimeError):
    pass


class Region(C.Structure):
    _fields_ = [("seed", C.c_int64), ("x0", C.c_int32), ("z0", C.c_int32), ("nx", C.c_int32), ("nz", C.c_int32),
                ("ny", C.c_int32), ("y", C.c_int32 * 8), ("zoom", C.c_int32), ("world", C.c_int32),
                ("out_off", C.c_int64)]


class McgHost(C.Structure):
    _fields_ = [("version", C.c_int), ("dim", C.c_int), ("mode", C.c_int), ("preset", C.c_char * 16),
                ("impl", C.c_void_p)]


_lib = None


class Obs(C.Structure):
    """One observation (include/mcbiome.h McbObs): point, flags (OBS_BLOCK, End radius), biome set."""
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32), ("flags", C.c_int32),
                ("mask", C.c_uint64 * 2)]


class SeedSource(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("list", C.POINTER(C.c_int64)), ("base", C.c_uint64),
                ("n", C.c_uint64)]


OBS_BLOCK = 0x100
SEARCH_TIES, SEARCH_NOPRUNE, SEARCH_SLOTS = 1, 2, 4


def make_obs(x, y, z, biomes, block=False, radius=12):
    """Observation accepting any of `biomes` (ids) at quart (x, y, z), or block coordinates if block."""
    o = Obs()
    o.x, o.y, o.z = x, y, z
    o.flags = (OBS_BLOCK if block else 0) | (radius & 0xFF)
    m = [0, 0]
    for b in biomes:
        m[b >> 6] |= 1 << (b & 63)
    o.mask[0], o.mask[1] = m
    return o


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EngineError("HIP engine not built: %s missing (run __graft_entry__.build())" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P, i64p, i32p, u8p = C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)
        L.mcb_last_error.restype = C.c_char_p
        L.mcb_open.argtypes = [C.POINTER(P), C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_int]
        L.mcb_close.argtypes = [P]
        L.mcb_engine_info.argtypes = [P, i32p, i32p, i32p, i32p]
        L.mcb_engine_cells.argtypes = [P, i64p, i64p, C.POINTER(C.c_double)]
        L.mcb_targets.argtypes = [P, i32p, C.c_int, u8p]
        L.mcb_search.argtypes = [P, C.POINTER(SeedSource), C.POINTER(Obs), C.c_int, C.c_int, i64p, C.c_uint32,
                                 C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
        L.mcb_end_obs_radius.argtypes = [P, C.POINTER(C.c_uint64)]
        L.mcb_plugin_world.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_void_p)]
        L.mcb_plugin_world_free.argtypes = [C.c_void_p]
        L.mcb_plugin_terrain_world.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_void_p)]
        L.mcb_plugin_terrain_world_free.argtypes = [C.c_void_p]
        L.mcb_terrain_batch_device.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                               C.POINTER(C.c_void_p)]
        L.mcb_chunk_biomes.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_uint8)]
        L.mcb_end_field.argtypes = [P, i64p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_double)]
        L.mcb_points.argtypes = [P, i64p, C.c_int, i32p, C.c_int, u8p, i32p]
        L.mcb_points_dev.argtypes = [P, P, C.c_int, P, C.c_int, P, P, P]
        L.mcb_points_dev_noalloc.argtypes = [P, P, C.c_int, P, C.c_int, P, P, P]
        L.mcb_regions_bytes.restype = C.c_int64
        L.mcb_regions_bytes.argtypes = [C.POINTER(Region), C.c_int]
        L.mcb_region.argtypes = [P, C.POINTER(Region), u8p]
        L.mcb_plan_create.argtypes = [P, C.POINTER(Region), C.c_int, C.c_int, C.c_int, C.POINTER(P)]
        L.mcb_plan_run.argtypes = [P, P, P]
        L.mcb_plan_run_graph.argtypes = [P, P, P]
        L.mcb_plan_run_stage.argtypes = [P, C.c_int, P, P]
        L.mcb_plan_info.argtypes = [P, i64p, i64p, i64p, i64p]
        L.mcb_plan_shape.argtypes = [P, C.c_int, i32p, i32p, i32p, i32p]
        L.mcb_plan_free.argtypes = [P]
        L.mcb_plan_autotune.argtypes = [P, P, P, i32p, C.POINTER(C.c_double)]
        L.mcb_context_create.argtypes = [C.POINTER(P), C.c_char_p]
        L.mcb_context_free.argtypes = [P]
        L.mcb_context_world.argtypes = [P, C.c_int, C.c_int, C.c_char_p]
        L.mcb_context_plan.argtypes = [P, C.POINTER(Region), C.c_int, C.c_int, C.c_int, C.POINTER(P)]
        L.mcb_server_create.argtypes = [C.POINTER(P), P, C.c_int64, C.c_int, C.c_int]
        L.mcb_server_submit.argtypes = [P, C.POINTER(Region), P, i64p]
        L.mcb_server_wait.argtypes = [P, C.c_int64, C.POINTER(C.c_double)]
        L.mcb_server_drain.argtypes = [P]
        L.mcb_server_stats.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), i64p,
                                       C.POINTER(C.c_double), i64p]
        L.mcb_server_reset_stats.argtypes = [P]
        L.mcb_server_slowest.argtypes = [P, C.c_int, C.POINTER(C.c_double)]
        L.mcb_server_destroy.argtypes = [P]
        L.mcg_host_load.argtypes = [C.POINTER(McgHost), C.c_int, C.c_int, C.c_char_p]
        L.mcg_host_free.argtypes = [C.POINTER(McgHost)]
        L.mcg_ref_points.argtypes = [C.POINTER(McgHost), i64p, C.c_int, i32p, C.c_int, u8p, i32p]
        L.mcg_ref_tie_mask.argtypes = [C.POINTER(McgHost), i32p, C.POINTER(C.c_uint64)]
        L.mcg_ref_set_error_handler.argtypes = [C.c_void_p, C.c_void_p]
        L.mcb_tie_masks.argtypes = [P, i32p, C.c_int, C.POINTER(C.c_uint64)]
        L.mcb_tree.argtypes = [C.c_int, C.c_int, C.c_char_p, i32p, C.c_int]
        L.mcb_climate_raw.argtypes = [P, i64p, C.c_int, i32p, C.c_int, C.POINTER(C.c_float)]
        L.mcb_noise_values.argtypes = [P, C.c_int, i64p, C.c_int, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double)]
        L.mcg_biome_name.restype = C.c_char_p
        L.mcg_biome_id.argtypes = [C.c_char_p]
        L.mcg_version_name.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise EngineError(lib().mcb_last_error().decode())


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def biome_names():
    L = lib()
    return [L.mcg_biome_name(i).decode() for i in range(L.mcg_biome_count())]


def device_count():
    return lib().mcb_device_count()


def tree_nodes(version, dim, data_dir=None):
    """mcb_tree: the engine's R-tree (built on the host from the parameter list), (n, 16) int32 rows
    lo0, hi0 .. lo5, hi5, offset distance^2, escape index, leaf biome (-1 = subtree), 0."""
    d = data_dir.encode() if data_dir else None
    n = lib().mcb_tree(VERSIONS[version], DIMS[dim], d, None, 0)
    if n < 0:
        raise EngineError(lib().mcb_last_error().decode())
    out = np.zeros(16 * n, np.int32)
    _check(0 if lib().mcb_tree(VERSIONS[version], DIMS[dim], d, _p(out, C.c_int32), out.size) == n else 1)
    return out.reshape(n, 16)


def make_region(seed, x0, z0, nx, nz, ys, zoom=False, out_off=0):
    r = Region()
    r.seed, r.x0, r.z0, r.nx, r.nz = int(seed), int(x0), int(z0), int(nx), int(nz)
    ys = list(ys)
    if len(ys) > 8:
        raise EngineError("at most 8 layers per region")
    r.ny = len(ys)
    for j, y in enumerate(ys):
        r.y[j] = int(y)
    r.zoom = 1 if zoom else 0
    r.out_off = int(out_off)
    return r


class Host:
    """mcg_host_load / mcg_ref_points drop-in (engine/cuda/mcg_host.h:34-47), on the HIP device."""

    def __init__(self, version="26.3", dim="overworld", preset="normal"):
        self.h = McgHost()
        if lib().mcg_host_load(C.byref(self.h), VERSIONS[version], DIMS[dim], preset.encode()) != 0:
            raise EngineError("mcg_host_load: " + lib().mcb_last_error().decode())

    @property
    def mode(self):
        return self.h.mode

    def ref_points(self, seeds, pts, want_tg=False):
        seeds = np.ascontiguousarray(seeds, dtype=np.int64)
        pts = np.ascontiguousarray(pts, dtype=np.int32).reshape(-1, 3)
        ns, np_ = seeds.size, pts.shape[0]
        out = np.zeros(ns * np_, np.uint8)
        tg = np.zeros(ns * np_ * 6, np.int32) if want_tg else None
        lib().mcg_ref_points(C.byref(self.h), _p(seeds, C.c_int64), ns, _p(pts, C.c_int32), np_, _p(out, C.c_uint8),
                             _p(tg, C.c_int32) if want_tg else None)
        out = out.reshape(ns, np_)
        return (out, tg.reshape(ns, np_, 6)) if want_tg else out

    def tie_mask(self, tg6):
        """mcg_ref_tie_mask: (mask lo, mask hi) of the biomes of every least-distance leaf."""
        t = np.ascontiguousarray(tg6, dtype=np.int32).reshape(6)
        m = (C.c_uint64 * 2)()
        lib().mcg_ref_tie_mask(C.byref(self.h), _p(t, C.c_int32), m)
        return m[0], m[1]

    def close(self):
        if self.h.impl:
            lib().mcg_host_free(C.byref(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    def __init__(self, engine, regions, tile_w=0, blocks=0, context=None):
        self.engine = engine
        self.context = context
        arr = (Region * len(regions))(*regions)
        self._arr = arr
        self.h = C.c_void_p()
        if context is not None:
            _check(lib().mcb_context_plan(context.h, arr, len(regions), tile_w, blocks, C.byref(self.h)))
        else:
            _check(lib().mcb_plan_create(engine.h, arr, len(regions), tile_w, blocks, C.byref(self.h)))
        self.out_bytes = lib().mcb_regions_bytes(arr, len(regions))
        t, c, s, n = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().mcb_plan_info(self.h, C.byref(t), C.byref(c), C.byref(s), C.byref(n)))
        self.tiles, self.columns, self.samples, self.seeds = t.value, c.value, s.value, n.value

    def run(self, d_out_ptr, stream_ptr=0, graph=False):
        fn = lib().mcb_plan_run_graph if graph else lib().mcb_plan_run
        _check(fn(self.h, C.c_void_p(d_out_ptr), C.c_void_p(stream_ptr or None)))

    def shape(self, group=0):
        """(tile_w, cols_per_lane, blocks, wrap_free) of a world group's launches."""
        v = [C.c_int32() for _ in range(4)]
        _check(lib().mcb_plan_shape(self.h, group, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune(self, d_out_ptr, stream_ptr=0):
        tw, ms = C.c_int32(), C.c_double()
        _check(lib().mcb_plan_autotune(self.h, C.c_void_p(d_out_ptr), C.c_void_p(stream_ptr or None), C.byref(tw),
                                       C.byref(ms)))
        return tw.value, ms.value

    def run_stage(self, stage, d_out_ptr, stream_ptr=0):
        _check(lib().mcb_plan_run_stage(self.h, stage, C.c_void_p(d_out_ptr), C.c_void_p(stream_ptr or None)))

    def free(self):
        if self.h:
            lib().mcb_plan_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Engine:
    """Native engine: one (version, dimension, preset) world on the current HIP device."""

    def __init__(self, version="26.3", dim="overworld", preset="normal", data_dir=None):
        self.version, self.dim, self.preset = version, dim, preset
        self.h = C.c_void_p()
        _check(lib().mcb_open(C.byref(self.h), VERSIONS[version], DIMS[dim], preset.encode(),
                              data_dir.encode() if data_dir else None, 0))
        v, d, m, n = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().mcb_engine_info(self.h, C.byref(v), C.byref(d), C.byref(m), C.byref(n)))
        self.mode, self.tree_nodes = m.value, n.value

    def cells(self):
        """(cells, list entries, build ms) of the Overworld climate cell table; zeros elsewhere."""
        c, l, ms = C.c_int64(), C.c_int64(), C.c_double()
        _check(lib().mcb_engine_cells(self.h, C.byref(c), C.byref(l), C.byref(ms)))
        return c.value, l.value, ms.value

    def targets(self, tg):
        """Biomes of quantized climate targets (n x {t,h,c,e,d,w})."""
        t = np.ascontiguousarray(tg, dtype=np.int32).reshape(-1, 6)
        out = np.zeros(t.shape[0], np.uint8)
        _check(lib().mcb_targets(self.h, _p(t, C.c_int32), t.shape[0], _p(out, C.c_uint8)))
        return out

    def climate_raw(self, seeds, xyz):
        """(ns, np, 6) float32 raw climate at block positions xyz (np, 3)."""
        seeds = np.ascontiguousarray(seeds, dtype=np.int64).ravel()
        xyz = np.ascontiguousarray(xyz, dtype=np.int32).reshape(-1, 3)
        out = np.zeros(seeds.size * len(xyz) * 6, np.float32)
        _check(lib().mcb_climate_raw(self.h, _p(seeds, C.c_int64), seeds.size, _p(xyz, C.c_int32), len(xyz),
                                     _p(out, C.c_float)))
        return out.reshape(seeds.size, len(xyz), 6)

    NOISES = {"overworld": ["offset", "temperature", "vegetation", "continentalness", "erosion", "ridge"],
              "the_nether": ["nether/temperature", "nether/vegetation"]}

    def noise_values(self, noise, seeds, xyz):
        """(ns, np) float64 NormalNoise values; noise = index or name (NOISES)."""
        if isinstance(noise, str):
            noise = self.NOISES["overworld" if self.dim == "overworld" else "the_nether"].index(noise)
        seeds = np.ascontiguousarray(seeds, dtype=np.int64).ravel()
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(seeds.size * len(xyz), np.float64)
        _check(lib().mcb_noise_values(self.h, int(noise), _p(seeds, C.c_int64), seeds.size, _p(xyz, C.c_double), len(xyz),
                                      _p(out, C.c_double)))
        return out.reshape(seeds.size, len(xyz))

    def tie_masks(self, tg):
        """(n, 2) uint64 biome masks of every least-distance leaf of each quantized target."""
        t = np.ascontiguousarray(tg, dtype=np.int32).reshape(-1, 6)
        out = np.zeros(2 * t.shape[0], np.uint64)
        _check(lib().mcb_tie_masks(self.h, _p(t, C.c_int32), t.shape[0], out.ctypes.data_as(C.POINTER(C.c_uint64))))
        return out.reshape(-1, 2)

    def search(self, obs, seeds=None, base=0, n=0, structural=False, flags=0, max_found=1 << 16):
        """Seed search: candidates = `seeds` (list), or base|i<<48 (structural), or base+i for i < n.
        Returns (found int64 array sorted as unsigned, reached uint64[len(obs)+1], device ms)."""
        arr = (Obs * len(obs))(*obs)
        src = SeedSource()
        keep = None
        if seeds is not None:
            keep = np.ascontiguousarray(seeds, dtype=np.int64)
            src.kind, src.list, src.n = 0, _p(keep, C.c_int64), keep.size
        elif structural:
            src.kind, src.base = 1, base & ((1 << 64) - 1)
        else:
            src.kind, src.base, src.n = 2, base & ((1 << 64) - 1), n
        found = np.zeros(max(1, max_found), np.int64)
        nf = C.c_uint32()
        reached = np.zeros(len(obs) + 1, np.uint64)
        ms = C.c_double()
        _check(lib().mcb_search(self.h, C.byref(src), arr, len(obs), flags, _p(found, C.c_int64), max_found,
                                C.byref(nf), reached.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(ms)))
        return found[: nf.value].copy(), reached, ms.value

    def end_obs_radius(self, biomes):
        m = (C.c_uint64 * 2)()
        for b in biomes:
            m[b >> 6] |= 1 << (b & 63)
        r = lib().mcb_end_obs_radius(self.h, m)
        if r < 0:
            raise EngineError(lib().mcb_last_error().decode())
        return r

    def chunk_biomes(self, seed, chunks):
        """Whole chunks (cx, cz) as libmcgen stores them: (n, qh, 4, 4) uint8 indexed [chunk, qy, qz, qx]."""
        cxz = np.ascontiguousarray(chunks, dtype=np.int32).reshape(-1, 2)
        qh = 96 if DIMS[self.dim] == 0 else 64
        out = np.empty(len(cxz) * qh * 16, np.uint8)
        _check(lib().mcb_chunk_biomes(self.h, int(seed), len(cxz), _p(cxz, C.c_int32), _p(out, C.c_uint8)))
        return out.reshape(len(cxz), qh, 4, 4)

    def plugin_world(self, seed):
        """This world for `seed` in libmcgen GPU-plugin form (a McgBiomeWorld, include/mcgpu.h): pass
        ``.addr`` to mcgpu_biome_new (mcgpu.BiomeWorld)."""
        return PluginWorld(self, seed)

    def terrain_world(self, seed):
        """The stand-in Overworld terrain world (mcb_plugin_terrain_world) of a 26.3+ Overworld engine for
        `seed`: ``.world`` feeds mcgpu_terrain_new (mcgpu.TerrainWorld)."""
        return PluginTerrainWorld(self, seed)

    def end_field(self, seeds, bx0, bz0, nx, nz, step=1, kind=1):
        """End island field: kind 0 height value at section (bx/8, bz/8), kind 1 erosion at the block;
        returns (n_seeds, nz, nx) float64."""
        seeds = np.ascontiguousarray(seeds, dtype=np.int64).ravel()
        out = np.zeros(seeds.size * nx * nz, np.float64)
        _check(lib().mcb_end_field(self.h, _p(seeds, C.c_int64), seeds.size, bx0, bz0, nx, nz, step, kind,
                                   _p(out, C.c_double)))
        return out.reshape(seeds.size, nz, nx)

    def points(self, seeds, pts, want_tg=False):
        seeds = np.ascontiguousarray(seeds, dtype=np.int64)
        pts = np.ascontiguousarray(pts, dtype=np.int32).reshape(-1, 3)
        ns, np_ = seeds.size, pts.shape[0]
        out = np.zeros(ns * np_, np.uint8)
        tg = np.zeros(ns * np_ * 6, np.int32) if want_tg else None
        _check(lib().mcb_points(self.h, _p(seeds, C.c_int64), ns, _p(pts, C.c_int32), np_, _p(out, C.c_uint8),
                                _p(tg, C.c_int32) if want_tg else None))
        out = out.reshape(ns, np_)
        return (out, tg.reshape(ns, np_, 6)) if want_tg else out

    def points_dev(self, d_seeds, ns, d_pts, np_, d_out, d_tg=0, stream=0):
        _check(lib().mcb_points_dev(self.h, C.c_void_p(d_seeds), ns, C.c_void_p(d_pts), np_, C.c_void_p(d_out),
                                    C.c_void_p(d_tg or None), C.c_void_p(stream or None)))

    def region(self, seed, x0, z0, nx, nz, ys, zoom=False):
        r = make_region(seed, x0, z0, nx, nz, ys, zoom)
        out = np.zeros(len(list(ys)) * nx * nz, np.uint8)
        _check(lib().mcb_region(self.h, C.byref(r), _p(out, C.c_uint8)))
        return out.reshape(len(list(ys)), nz, nx)

    def plan(self, regions, tile_w=0, blocks=0):
        return Plan(self, regions, tile_w, blocks)

    def close(self):
        if self.h:
            lib().mcb_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """Several worlds on one device (mixed batches, serving)."""

    def __init__(self, data_dir=None):
        self.h = C.c_void_p()
        _check(lib().mcb_context_create(C.byref(self.h), data_dir.encode() if data_dir else None))

    def world(self, version, dim, preset="normal"):
        w = lib().mcb_context_world(self.h, VERSIONS[version], DIMS[dim], preset.encode())
        if w < 0:
            raise EngineError(lib().mcb_last_error().decode())
        return w

    def plan(self, regions, tile_w=0, blocks=0):
        return Plan(None, regions, tile_w, blocks, context=self)

    def close(self):
        if self.h:
            lib().mcb_context_free(self.h)
            self.h = None


class Server:
    """Continuous-batching request server over a Context (mcb_server_*)."""

    def __init__(self, context, batch_bytes=8 << 20, lanes=4, slot_capacity=16384):
        self.context = context
        self.h = C.c_void_p()
        _check(lib().mcb_server_create(C.byref(self.h), context.h, batch_bytes, lanes, slot_capacity))

    def submit(self, region, host_out):
        """host_out: a writable contiguous uint8 numpy array of ny*nx*nz bytes (kept alive by the caller)."""
        t = C.c_int64()
        _check(lib().mcb_server_submit(self.h, C.byref(region), host_out.ctypes.data_as(C.c_void_p), C.byref(t)))
        return t.value

    def wait(self, ticket):
        ms = C.c_double()
        _check(lib().mcb_server_wait(self.h, ticket, C.byref(ms)))
        return ms.value

    def drain(self):
        _check(lib().mcb_server_drain(self.h))

    def stats(self):
        p50, p99, p999, hit = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        n, b = C.c_int64(), C.c_int64()
        _check(lib().mcb_server_stats(self.h, C.byref(p50), C.byref(p99), C.byref(p999), C.byref(n), C.byref(hit),
                                      C.byref(b)))
        return {"p50_ms": p50.value, "p99_ms": p99.value, "p999_ms": p999.value, "completed": n.value,
                "slot_hit_rate": hit.value, "batches": b.value}

    def slowest(self, n=5):
        """The n slowest requests: [(total, queued, launch, device, copy) ms], slowest first."""
        out = np.zeros(5 * n, np.float64)
        k = lib().mcb_server_slowest(self.h, n, out.ctypes.data_as(C.POINTER(C.c_double)))
        _check(0 if k >= 0 else k)
        return [tuple(round(float(v), 3) for v in out[5 * i: 5 * i + 5]) for i in range(k)]

    def reset_stats(self):
        _check(lib().mcb_server_reset_stats(self.h))

    def close(self):
        if self.h:
            lib().mcb_server_destroy(self.h)
            self.h = None
