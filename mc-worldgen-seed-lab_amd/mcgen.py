"""ctypes binding of the libmcgen-compatible biome API exported by libmcbiome_hip.so (include/mcgen.h).

The signatures and structures are libmcgen's (libmcgen/include/mcgen.h), so a Python caller written for
libmcgen (e.g. blender/mcgen_addon/core/lib.py:100-140 binds the same names) can load this library instead
for the biome stage.
"""
import ctypes as C

import numpy as np

import mcbiome

MCGEN_OK, MCGEN_E_ARG, MCGEN_E_VERSION, MCGEN_E_CANCEL, MCGEN_E_UNSUPPORTED = 0, 1, 4, 6, 7
MC_STAGE_BIOMES, MC_STAGE_TERRAIN = 1, 2
MCGEN_COMPUTE_CPU, MCGEN_COMPUTE_GPU, MCGEN_COMPUTE_AUTO = 0, 1, 2


class McSeeds(C.Structure):
    _fields_ = [("climate", C.c_int64), ("terrain", C.c_int64), ("structures", C.c_int64), ("features", C.c_int64)]


class McTweakValue(C.Structure):
    _fields_ = [("id", C.c_char_p), ("value", C.c_double)]


class McRegionInfo(C.Structure):
    _fields_ = [("cx0", C.c_int), ("cz0", C.c_int), ("nx", C.c_int), ("nz", C.c_int), ("min_y", C.c_int),
                ("height", C.c_int)]


PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_char_p)
_bound = None


def lib():
    global _bound
    L = mcbiome.lib()
    if _bound is not L:
        P, u8p, ip = C.c_void_p, C.POINTER(C.c_uint8), C.POINTER(C.c_int)
        sig = {
            "mcgen_version": (C.c_char_p, []), "mcgen_abi_version": (C.c_int, []),
            "mcgen_open": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(P), C.c_char_p, C.c_size_t]),
            "mcgen_close": (None, [P]),
            "mcgen_dimension_count": (C.c_int, [P]), "mcgen_dimension_name": (C.c_char_p, [P, C.c_int]),
            "mcgen_preset_count": (C.c_int, [P, C.c_char_p]), "mcgen_preset_name": (C.c_char_p, [P, C.c_char_p, C.c_int]),
            "mcgen_block_state_count": (C.c_int, [P]), "mcgen_biome_count": (C.c_int, [P]),
            "mcgen_biome_name": (C.c_char_p, [P, C.c_int]), "mcgen_tweak_count": (C.c_int, [P]),
            "mcgen_world_new": (C.c_int, [P, C.c_char_p, C.c_char_p, C.POINTER(McSeeds), C.POINTER(McTweakValue), C.c_int,
                                          C.POINTER(P), C.c_char_p, C.c_size_t]),
            "mcgen_world_free": (None, [P]),
            "mcgen_world_min_y": (C.c_int, [P]), "mcgen_world_height": (C.c_int, [P]),
            "mcgen_world_sea_level": (C.c_int, [P]),
            "mcgen_biome_at": (C.c_int, [P, C.c_int, C.c_int, C.c_int]),
            "mcgen_biome_grid": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, u8p]),
            "mcgen_biome_grid_cpu": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, u8p]),
            "mcgen_gpu_biome_grid": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, u8p]),
            "mcgen_gpu_biome_points": (C.c_int, [P, C.c_int, ip, u8p, ip]),
            "mcgen_generate_region": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int, PROGRESS_FN,
                                                P, C.POINTER(P), C.c_char_p, C.c_size_t]),
            "mcgen_region_free": (None, [P]), "mcgen_region_info": (None, [P, C.POINTER(McRegionInfo)]),
            "mcgen_region_biomes": (P, [P, C.c_int, C.c_int]), "mcgen_region_blocks": (P, [P, C.c_int, C.c_int]),
            "mcgen_gpu_set_compute": (C.c_int, [C.c_int, C.c_int]), "mcgen_gpu_get_compute": (C.c_int, []),
            "mcgen_gpu_status": (C.c_int, [C.c_char_p, C.c_size_t]),
            "mcgen_gpu_selftest": (C.c_int, [P, C.c_int, C.c_char_p, C.c_size_t]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _bound = L
    return L


class McgenError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%d: %s" % (code, msg))
        self.code = code


class Gen:
    def __init__(self, version, pack_dir=None):
        self.h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib().mcgen_open(pack_dir.encode() if pack_dir else None, version.encode(), C.byref(self.h), err, 512)
        if rc:
            raise McgenError(rc, err.value.decode())

    def world(self, dimension, preset, seed, tweaks=()):
        return World(self, dimension, preset, seed, tweaks)

    def biome_names(self):
        """The version's biome registry in id order ("minecraft:<name>")."""
        L = lib()
        return [L.mcgen_biome_name(self.h, i).decode() for i in range(L.mcgen_biome_count(self.h))]

    def close(self):
        if self.h:
            lib().mcgen_close(self.h)
            self.h = None


class World:
    def __init__(self, gen, dimension, preset, seed, tweaks=()):
        self.h = C.c_void_p()
        s = McSeeds(seed, seed, seed, seed)
        tw = (McTweakValue * max(1, len(tweaks)))(*[McTweakValue(k.encode(), v) for k, v in tweaks])
        err = C.create_string_buffer(512)
        rc = lib().mcgen_world_new(gen.h, dimension.encode(), preset.encode() if preset else None, C.byref(s), tw,
                                   len(tweaks), C.byref(self.h), err, 512)
        if rc:
            raise McgenError(rc, err.value.decode())
        self._gen = gen

    def biome_at(self, x, y, z):
        return lib().mcgen_biome_at(self.h, x, y, z)

    def biome_grid(self, x0, z0, nx, nz, step, y):
        out = np.empty(nx * nz, np.uint8)
        rc = lib().mcgen_biome_grid(self.h, x0, z0, nx, nz, step, y, out.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc:
            raise McgenError(rc, "mcgen_biome_grid")
        return out.reshape(nz, nx)

    def gpu_biome_grid(self, x0, z0, nx, nz, step, y, out=None):
        """mcgen_gpu_biome_grid into `out` (a uint8 array of nx*nz, allocated when None)."""
        if out is None:
            out = np.empty(nx * nz, np.uint8)
        rc = lib().mcgen_gpu_biome_grid(self.h, x0, z0, nx, nz, step, y, out.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc:
            raise McgenError(rc, "mcgen_gpu_biome_grid")
        return out.reshape(nz, nx)

    def region_biomes(self, cx0, cz0, nx, nz):
        """MC_STAGE_BIOMES region: (nz, nx, height/4, 4, 4) [chunk z, chunk x, qy, qz, qx]."""
        r = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib().mcgen_generate_region(self.h, cx0, cz0, nx, nz, MC_STAGE_BIOMES, 0, PROGRESS_FN(0), None,
                                         C.byref(r), err, 512)
        if rc:
            raise McgenError(rc, err.value.decode())
        qh = lib().mcgen_world_height(self.h) // 4
        out = np.empty((nz, nx, qh, 4, 4), np.uint8)
        for iz in range(nz):
            for ix in range(nx):
                p = lib().mcgen_region_biomes(r, cx0 + ix, cz0 + iz)
                C.memmove(out[iz, ix].ctypes.data, p, qh * 16)
        lib().mcgen_region_free(r)
        return out

    def close(self):
        if self.h:
            lib().mcgen_world_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
