"""ctypes view of the libmcgen GPU plugin ABI exported by libmcbiome_hip.so (include/mcgpu.h).

libmcgen itself binds these symbols from C (libmcgen/src/gpu_bridge.c:27-148); this module is the same
binding for Python callers and tests: the ABI structures as ctypes.Structure (field for field the
layout of libmcgen/gpu/mcgen_gpu_abi.h) and a thin handle class over mcgpu_biome_new / _grid / _points.
"""
import ctypes as C

import numpy as np

import mcbiome

MCGPU_ABI = 2
MCG_MAX_ROOTS = 16
MCG_BS_MULTI, MCG_BS_END, MCG_BS_FIXED = 0, 1, 2

# node kinds of 26.3+ (float) programs
K = dict((n, i) for i, n in enumerate(
    "CONST NOISE NOISE_XZ NOISE_XYZ SHIFT_B END DIST GRAD_CLAMP GRAD_REPEAT GRAD_MIRROR CTX_ALPHA CTX_OFFSET "
    "CTX_BEARD ABS SQUARE CUBE SQRT LEAKY RECIP NEG SQUEEZE LOG SIGN ROUND_INT ROUND ADD CADD CSUB SUB MUL CMUL "
    "DIV CDIV MIN CMIN MAX CMAX POW_CB POW_CE POW SPLINE LERP LERP_CF LERP_CS CLAMP RANGE_C RANGE ISEL1 ISEL "
    "CACHE BLEND_DENSITY INTERP SLICE_X SLICE_Y SLICE_Z SLICE_XZ FTS".split()))
# node kinds of 26.1/26.2 (double) programs
O = dict((n, i) for i, n in enumerate(
    "CONST NOISE SHIFTED_NOISE SHIFT_A SHIFT_B SHIFT END_ISLANDS Y_GRADIENT WEIRD_SCALED BLEND_ALPHA ZERO ABS "
    "SQUARE CUBE HALF_NEG QUARTER_NEG RECIP SQUEEZE CLAMP ADD MUL MIN MAX CADD CMUL RANGE ISEL SPLINE PASS".split()))


class McgOct(C.Structure):
    _fields_ = [("xo", C.c_double), ("yo", C.c_double), ("zo", C.c_double), ("p", C.c_uint8 * 256)]


class McgLayer(C.Structure):
    _fields_ = [("a", C.c_double), ("b", C.c_double), ("amp_d", C.c_double), ("amp_f", C.c_float), ("oct", C.c_int),
                ("kind", C.c_int), ("pad", C.c_int)]


class McgNoise(C.Structure):
    _fields_ = [("l0", C.c_int), ("n0", C.c_int), ("l1", C.c_int), ("n1", C.c_int), ("value_factor", C.c_double)]


class McgSpline(C.Structure):
    _fields_ = [("is_const", C.c_int), ("value", C.c_float), ("coord", C.c_int), ("n", C.c_int), ("loc0", C.c_int),
                ("val0", C.c_int)]


class McgNode(C.Structure):
    _fields_ = [("k", C.c_int), ("a", C.c_int), ("b", C.c_int), ("c", C.c_int), ("ns", C.c_int), ("sp", C.c_int),
                ("arr0", C.c_int), ("narr", C.c_int), ("thr0", C.c_int), ("nthr", C.c_int),
                ("i0", C.c_int), ("i1", C.c_int), ("i2", C.c_int), ("i3", C.c_int),
                ("f0", C.c_float), ("f1", C.c_float), ("f2", C.c_float), ("f3", C.c_float),
                ("d0", C.c_double), ("d1", C.c_double),
                ("cm", C.c_int), ("cx", C.c_int), ("cy", C.c_int), ("cz", C.c_int), ("oct", C.c_int), ("cid", C.c_int)]


class McgProg(C.Structure):
    _fields_ = [("old", C.c_int),
                ("nnodes", C.c_int), ("nodes", C.POINTER(McgNode)),
                ("nnoise", C.c_int), ("noise", C.POINTER(McgNoise)),
                ("nlayers", C.c_int), ("layers", C.POINTER(McgLayer)),
                ("noct", C.c_int), ("octs", C.POINTER(McgOct)),
                ("nsp", C.c_int), ("sp", C.POINTER(McgSpline)),
                ("nsplf", C.c_int), ("splf", C.POINTER(C.c_float)),
                ("nspc", C.c_int), ("spc", C.POINTER(C.c_int)),
                ("narr", C.c_int), ("arr", C.POINTER(C.c_int)),
                ("nthr", C.c_int), ("thr", C.POINTER(C.c_double)),
                ("nroots", C.c_int), ("root", C.c_int * MCG_MAX_ROOTS),
                ("ncache", C.c_int)]


class McgTreeNode(C.Structure):
    _fields_ = [("lo", C.c_int32 * 7), ("hi", C.c_int32 * 7), ("first", C.c_int32), ("count", C.c_int32),
                ("biome", C.c_int32), ("pad", C.c_int32)]


MCG_MAX_VEINS = 4


class McgVein(C.Structure):
    _fields_ = [("density_root", C.c_int), ("richness_root", C.c_int), ("gap_root", C.c_int), ("ore", C.c_int),
                ("raw", C.c_int), ("filler", C.c_int), ("raw_chance", C.c_float)]


class McgTerrainWorld(C.Structure):
    _fields_ = [("vol", C.POINTER(McgProg)), ("gap", C.POINTER(McgProg)), ("nmin", C.c_int), ("nh", C.c_int),
                ("root_density", C.c_int), ("nveins", C.c_int), ("vein", McgVein * MCG_MAX_VEINS),
                ("ore_lo", C.c_uint64), ("ore_hi", C.c_uint64), ("veins_on", C.c_int),
                ("cache_ids", C.c_int * 32), ("ncache_ids", C.c_int), ("pre_root", C.c_int),
                ("pre_sx", C.c_int), ("pre_sy", C.c_int), ("pre_sz", C.c_int), ("pre_dx", C.c_int), ("pre_dy", C.c_int),
                ("pre_dz", C.c_int), ("pre_y0", C.c_int), ("pre_xoff", C.c_int), ("pre_zoff", C.c_int)]


# Beardifier callback: (ud, chunk_index, vol[9], out) -> 0 when the chunk has no Beardifier
McgBeardFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float))


class McgBiomeWorld(C.Structure):
    _fields_ = [("source", C.c_int), ("fixed_biome", C.c_int), ("end_ids", C.c_int * 5), ("newf", C.c_int),
                ("min_qy", C.c_int), ("qh", C.c_int), ("zoom_seed", C.c_int64), ("prog", C.POINTER(McgProg)),
                ("ntree", C.c_int), ("tree", C.POINTER(McgTreeNode)), ("tree_root", C.c_int)]


_bound = None


def lib():
    """libmcbiome_hip.so with the mcgpu_* signatures bound (fails loudly when the library is missing)."""
    global _bound
    L = mcbiome.lib()
    if _bound is not L:
        P, ip, u8p = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_uint8)
        L.mcgpu_abi.restype = C.c_int
        L.mcgpu_device_count.restype = C.c_int
        L.mcgpu_device_info.argtypes = [C.c_int, C.c_char_p, C.c_size_t, ip, ip, C.POINTER(C.c_size_t)]
        L.mcgpu_select.argtypes = [C.c_int]
        L.mcgpu_selected.restype = C.c_int
        L.mcgpu_last_error.restype = C.c_char_p
        L.mcgpu_mem_info.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.mcgpu_biome_new.restype = P
        L.mcgpu_biome_new.argtypes = [P, C.c_char_p, C.c_size_t]
        L.mcgpu_biome_free.argtypes = [P]
        L.mcgpu_biome_grid.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, u8p, ip]
        L.mcgpu_biome_points.argtypes = [P, C.c_int, ip, u8p, ip]
        L.mcb_plugin_router.argtypes = [P]
        L.mcb_plugin_router_world.argtypes = [P]
        sp = C.POINTER(C.c_size_t)
        L.mcgpu_terrain_new.restype = P
        L.mcgpu_terrain_new.argtypes = [P, C.c_int, C.c_char_p, C.c_size_t]
        L.mcgpu_terrain_free.argtypes = [P]
        L.mcgpu_terrain_max_chunks.argtypes = [P]
        L.mcgpu_terrain_cell_floats.restype = C.c_size_t
        L.mcgpu_terrain_cell_floats.argtypes = [P]
        L.mcgpu_terrain_cell_info.argtypes = [P, C.c_int, ip, ip, sp, sp]
        L.mcgpu_terrain_cell_cand.argtypes = [P, C.c_int, C.c_int, ip, ip, ip, ip]
        L.mcgpu_terrain_batch.argtypes = [P, C.c_int, ip, ip, McgBeardFn, P, C.POINTER(C.c_float), C.POINTER(C.c_uint16),
                                          C.POINTER(C.c_float), C.POINTER(C.c_int8)]
        L.mcgpu_host_alloc.restype = P
        L.mcgpu_host_alloc.argtypes = [C.c_size_t]
        L.mcgpu_host_free.argtypes = [P]
        _bound = L
    return L


def last_error():
    return lib().mcgpu_last_error().decode(errors="replace")


class PluginError(RuntimeError):
    pass


def router_world(world):
    """Host-only check: would mcgpu_biome_new run this McgBiomeWorld (ctypes instance or address) through
    the engine's column code?"""
    addr = world if isinstance(world, int) else C.addressof(world)
    return bool(lib().mcb_plugin_router_world(C.c_void_p(addr)))


class BiomeWorld:
    """A world created with mcgpu_biome_new from a McgBiomeWorld (a ctypes instance or its address)."""

    def __init__(self, world):
        L = lib()
        addr = world if isinstance(world, int) else C.addressof(world)
        err = C.create_string_buffer(512)
        self.h = L.mcgpu_biome_new(C.c_void_p(addr), err, len(err))
        if not self.h:
            raise PluginError(err.value.decode(errors="replace"))

    def grid(self, x0, z0, nx, nz, step, y):
        out = np.empty(nx * nz, np.uint8)
        bad = C.c_int(0)
        rc = lib().mcgpu_biome_grid(self.h, x0, z0, nx, nz, step, y, out.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(bad))
        if rc:
            raise PluginError(last_error())
        return out.reshape(nz, nx), bad.value

    @property
    def router(self):
        """True when the world's program runs through the engine's fused column code (vanilla router)."""
        return bool(lib().mcb_plugin_router(self.h))

    def points(self, xyz):
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        out = np.empty(len(xyz), np.uint8)
        bad = C.c_int(0)
        rc = lib().mcgpu_biome_points(self.h, len(xyz), xyz.ctypes.data_as(C.POINTER(C.c_int)),
                                      out.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(bad))
        if rc:
            raise PluginError(last_error())
        return out, bad.value

    def close(self):
        if getattr(self, "h", None):
            lib().mcgpu_biome_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TerrainWorld:
    """A terrain world created with mcgpu_terrain_new from a McgTerrainWorld (a ctypes instance)."""

    def __init__(self, world, max_chunks=64):
        L = lib()
        err = C.create_string_buffer(512)
        self.world = world
        self.h = L.mcgpu_terrain_new(C.byref(world), max_chunks, err, len(err))
        if not self.h:
            raise PluginError(err.value.decode(errors="replace"))
        self.max_chunks = L.mcgpu_terrain_max_chunks(self.h)
        self.cell_floats = L.mcgpu_terrain_cell_floats(self.h)

    def cells(self):
        """[(cid, [(vol9, xrel, zrel, n), ...], cap, off)] per McgTerrainWorld.cache_ids entry"""
        L = lib()
        out = []
        for k in range(self.world.ncache_ids):
            cid, nc, cap, off = C.c_int(), C.c_int(), C.c_size_t(), C.c_size_t()
            if L.mcgpu_terrain_cell_info(self.h, k, C.byref(cid), C.byref(nc), C.byref(cap), C.byref(off)):
                raise PluginError("cell_info %d" % k)
            cands = []
            for j in range(nc.value):
                vol = (C.c_int * 9)()
                xr, zr, n = C.c_int(), C.c_int(), C.c_int()
                L.mcgpu_terrain_cell_cand(self.h, k, j, vol, C.byref(xr), C.byref(zr), C.byref(n))
                cands.append((tuple(vol), xr.value, zr.value, n.value))
            out.append((cid.value, cands, cap.value, off.value))
        return out

    def batch(self, cx, cz, beard=None, veins=True, cells=False):
        """(density (n, 16*nh*16) float32, veins (n, 16*nh*16) uint16 or None, cells (n, cell_floats) or None,
        which (n, ncache) int8 or None)"""
        cx = np.ascontiguousarray(cx, np.int32)
        cz = np.ascontiguousarray(cz, np.int32)
        n = len(cx)
        nd = 256 * self.world.nh
        dens = np.empty((n, nd), np.float32)
        vb = np.empty((n, nd), np.uint16) if veins and self.world.veins_on else None
        nc = self.world.ncache_ids
        cb = np.empty((n, max(self.cell_floats, 1)), np.float32) if cells else None
        wb = np.empty((n, max(nc, 1)), np.int8) if cells else None
        fn = beard if beard is not None else McgBeardFn()
        rc = lib().mcgpu_terrain_batch(self.h, n, cx.ctypes.data_as(C.POINTER(C.c_int)), cz.ctypes.data_as(C.POINTER(C.c_int)),
                                       fn, None, dens.ctypes.data_as(C.POINTER(C.c_float)),
                                       vb.ctypes.data_as(C.POINTER(C.c_uint16)) if vb is not None else None,
                                       cb.ctypes.data_as(C.POINTER(C.c_float)) if cb is not None else None,
                                       wb.ctypes.data_as(C.POINTER(C.c_int8)) if wb is not None else None)
        if rc:
            raise PluginError(last_error())
        return dens, vb, cb, wb

    def batch_device(self, cx, cz):
        """The density of the chunks left in device memory (mcb_terrain_batch_device): its device address,
        chunk i's 16*nh*16 floats at address + 4 * i * 16*nh*16, valid until the next batch."""
        cx = np.ascontiguousarray(cx, np.int32)
        cz = np.ascontiguousarray(cz, np.int32)
        p = C.c_void_p()
        rc = mcbiome.lib().mcb_terrain_batch_device(self.h, len(cx), cx.ctypes.data_as(C.POINTER(C.c_int)),
                                                    cz.ctypes.data_as(C.POINTER(C.c_int)), C.byref(p))
        if rc:
            raise PluginError(last_error())
        return p.value

    def close(self):
        if getattr(self, "h", None):
            lib().mcgpu_terrain_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
