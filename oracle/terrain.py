"""Terrain density checkers (test infrastructure only): ctypes bindings of

  * ``RefTerrain`` -- libmcgen's own density-function pipeline (parser, compiler, sampleVolume, GPU-program
    exporter; libmcgen/src/df_new.c and friends) behind oracle/ref_terrain.c, compiled from the reference's
    sources by ``make -C oracle ref`` into oracle/_ref/libmcgen_ref.so;
  * ``volume`` / ``chunk`` -- the CPU restatement in oracle/mcgterrain.c, over an exported program.

Worlds are described as the game's datapack JSON: a seed, noise parameters by id and density functions by
id (see tests/terrain_worlds.py).
"""
import ctypes as C
import json
import os

import numpy as np

import oracle as O

REF_TERRAIN_PATH = os.path.join(O.HERE, "_ref", "libmcgen_ref.so")


def _mcgpu():
    import mcgpu
    return mcgpu


_tlib = None


def _lib():
    global _tlib
    if _tlib is None:
        L = O.lib()
        M = _mcgpu()
        P = C.c_void_p
        ip, fp = C.POINTER(C.c_int), C.POINTER(C.c_float)
        L.or_terrain_volume.argtypes = [C.POINTER(M.McgProg), C.c_int, ip, fp]
        L.or_terrain_chunk.argtypes = [C.POINTER(M.McgTerrainWorld), C.c_int, C.c_int, C.c_int, M.McgBeardFn, P, fp,
                                       C.POINTER(C.c_uint16), ip, ip, ip, fp, C.c_int]
        L.or_terrain_ore_floats.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, fp]
        _tlib = L
    return _tlib


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def vol_size(v):
    return v[0] * v[1] * v[2]


def volume(prog, node, vol):
    """The restatement's sampleVolume of program node `node` over vol = (sx, sy, sz, x0, y0, z0, dx, dy, dz)."""
    M = _mcgpu()
    v = np.ascontiguousarray(vol, np.int32)
    out = np.zeros(vol_size(vol), np.float32)
    rc = _lib().or_terrain_volume(C.cast(prog, C.POINTER(M.McgProg)), node, _ip(v), _fp(out))
    if rc:
        raise RuntimeError("restatement: node kind or geometry not covered")
    return out


def chunk(world, cx, cz, index=0, beard=None, cap=4096):
    """One chunk of a McgTerrainWorld (ctypes instance) through the restatement: (density, veins or None,
    cells) with cells[k] = (set, changed, vol9, values) per world.cache_ids[k]."""
    M = _mcgpu()
    nd = 16 * world.nh * 16
    dens = np.zeros(nd, np.float32)
    veins = np.zeros(nd, np.uint16) if world.veins_on else None
    nc = world.ncache_ids
    st, ch = np.zeros(max(nc, 1), np.int32), np.zeros(max(nc, 1), np.int32)
    cvol = np.zeros(9 * max(nc, 1), np.int32)
    cvals = np.zeros(cap * max(nc, 1), np.float32)
    fn = beard if beard is not None else M.McgBeardFn()
    rc = _lib().or_terrain_chunk(C.byref(world), cx, cz, index, fn, None, _fp(dens),
                                 veins.ctypes.data_as(C.POINTER(C.c_uint16)) if veins is not None else None,
                                 _ip(st), _ip(ch), _ip(cvol), _fp(cvals), cap)
    if rc:
        raise RuntimeError("restatement: node kind or geometry not covered")
    cells = []
    for k in range(nc):
        v9 = tuple(int(t) for t in cvol[9 * k: 9 * k + 9])
        n = vol_size(v9) if st[k] else 0
        cells.append((int(st[k]), int(ch[k]), v9, cvals[k * cap: k * cap + min(n, cap)].copy()))
    return dens, veins, cells


def ore_floats(lo, hi, x, y, z, k=3):
    out = np.zeros(k, np.float32)
    _lib().or_terrain_ore_floats(lo, hi, x, y, z, k, _fp(out))
    return out


class RefTerrain:
    """A world of libmcgen's own density-function pipeline (oracle/_ref/libmcgen_ref.so)."""

    _L = None

    @classmethod
    def available(cls):
        return os.path.exists(REF_TERRAIN_PATH)

    @classmethod
    def _lib(cls):
        if cls._L is None:
            M = _mcgpu()
            L = C.CDLL(REF_TERRAIN_PATH)
            P, ip, fp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float)
            L.rt_open.restype = P
            L.rt_open.argtypes = [C.c_int64, C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]
            L.rt_close.argtypes = [P]
            L.rt_compile.argtypes = [P, C.c_char_p, C.c_char_p, C.c_size_t]
            L.rt_export.restype = C.POINTER(M.McgProg)
            L.rt_export.argtypes = [P, C.c_int, ip, C.c_int]
            L.rt_cache_ids.argtypes = [P, C.c_int, ip, C.c_int, ip]
            L.rt_chunk.argtypes = [P, C.c_int, ip, C.c_int, ip, fp, C.c_int, ip, ip, ip, ip, fp, C.c_int]
            L.rt_volume.argtypes = [P, C.c_int, ip, C.c_int, fp]
            L.rt_values.argtypes = [P, C.c_int, C.c_int, ip, fp]
            L.rt_ore.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
            L.rt_ore_floats.argtypes = [P, C.c_int, ip, C.c_int, fp]
            cls._L = L
        return cls._L

    def __init__(self, seed, noises, functions):
        L = self._lib()
        err = C.create_string_buffer(512)
        self.h = L.rt_open(seed, json.dumps(noises).encode(), json.dumps(functions).encode(), err, len(err))
        if not self.h:
            raise RuntimeError("libmcgen world: " + err.value.decode(errors="replace"))
        self.progs = {}

    def compile(self, df):
        """Compile a density function (dict, number or id) as the next root; returns its index."""
        err = C.create_string_buffer(512)
        r = self._lib().rt_compile(self.h, json.dumps(df).encode(), err, len(err))
        if r < 0:
            raise RuntimeError("libmcgen compile: " + err.value.decode(errors="replace"))
        return r

    def export(self, roots, point=False):
        """McgProg pointer of the roots (volume program; point=True: the point program with slices baked in).
        Valid until the next export of the same kind."""
        a = np.ascontiguousarray(roots, np.int32)
        p = self._lib().rt_export(self.h, 1 if point else 0, _ip(a), len(a))
        if not p:
            raise RuntimeError("libmcgen export failed")
        self.progs[point] = p
        return p

    def cache_ids(self, root):
        ids = np.zeros(32, np.int32)
        n = C.c_int(0)
        self._lib().rt_cache_ids(self.h, root, _ip(ids), 32, C.byref(n))
        return [int(t) for t in ids[:n.value]]

    def volume(self, root, vol, caches=True):
        v = np.ascontiguousarray(vol, np.int32)
        out = np.zeros(vol_size(vol), np.float32)
        self._lib().rt_volume(self.h, root, _ip(v), 1 if caches else 0, _fp(out))
        return out

    def chunk(self, dens, vol, pre=-1, pre_vol=None, cids=(), cap=4096):
        """One chunk as libmcgen's CPU terrain samples it: (density, cells) with cells as in ``chunk``."""
        v = np.ascontiguousarray(vol, np.int32)
        pv = np.ascontiguousarray(pre_vol if pre_vol is not None else [1] * 9, np.int32)
        out = np.zeros(vol_size(vol), np.float32)
        nc = len(cids)
        ci = np.ascontiguousarray(list(cids) or [0], np.int32)
        st, ch = np.zeros(max(nc, 1), np.int32), np.zeros(max(nc, 1), np.int32)
        cvol = np.zeros(9 * max(nc, 1), np.int32)
        cvals = np.zeros(cap * max(nc, 1), np.float32)
        self._lib().rt_chunk(self.h, pre, _ip(pv), dens, _ip(v), _fp(out), nc, _ip(ci), _ip(st), _ip(ch), _ip(cvol),
                             _fp(cvals), cap)
        cells = []
        for k in range(nc):
            v9 = tuple(int(t) for t in cvol[9 * k: 9 * k + 9])
            n = vol_size(v9) if st[k] else 0
            cells.append((int(st[k]), int(ch[k]), v9, cvals[k * cap: k * cap + min(n, cap)].copy()))
        return out, cells

    def values(self, root, xyz):
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        out = np.zeros(len(xyz), np.float32)
        self._lib().rt_values(self.h, root, len(xyz), _ip(xyz), _fp(out))
        return out

    def ore(self):
        lo, hi = C.c_uint64(0), C.c_uint64(0)
        self._lib().rt_ore(self.h, C.byref(lo), C.byref(hi))
        return lo.value, hi.value

    def ore_floats(self, xyz, k=3):
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        out = np.zeros(len(xyz) * k, np.float32)
        self._lib().rt_ore_floats(self.h, len(xyz), _ip(xyz), k, _fp(out))
        return out.reshape(len(xyz), k)

    def close(self):
        if getattr(self, "h", None):
            self._lib().rt_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def terrain_world(ref, density, nmin=-64, nh=384, veins=(), pre=-1, pre_geom=(11, 1, 11, 4, 1, 4, 0, -16, -16),
                  cache_ids=()):
    """A McgTerrainWorld over programs exported from ``ref`` (a RefTerrain): the volume program's roots are
    [density, vein density / richness pairs..., pre]; veins = [(density, richness, gap, ore, raw, filler,
    raw_chance)] with ref root indices; the gap roots go to the point program.  The programs stay valid
    until ``ref`` exports again."""
    M = _mcgpu()
    roots = [density]
    for v in veins:
        roots += [v[0], v[1]]
    pre_index = -1
    if pre >= 0:
        pre_index = len(roots)
        roots.append(pre)
    T = M.McgTerrainWorld()
    T.vol = ref.export(roots)
    if veins:
        T.gap = ref.export([v[2] for v in veins], point=True)
    T.nmin, T.nh, T.root_density = nmin, nh, 0
    T.nveins = len(veins)
    for k, v in enumerate(veins):
        T.vein[k] = M.McgVein(1 + 2 * k, 2 + 2 * k, k, v[3], v[4], v[5], v[6])
    T.ore_lo, T.ore_hi = ref.ore()
    T.veins_on = 1 if veins else 0
    for k, c in enumerate(cache_ids):
        T.cache_ids[k] = c
    T.ncache_ids = len(cache_ids)
    T.pre_root = pre_index
    (T.pre_sx, T.pre_sy, T.pre_sz, T.pre_dx, T.pre_dy, T.pre_dz, T.pre_y0, T.pre_xoff, T.pre_zoff) = pre_geom
    return T


def ref_chunk(ref, T, density, cx, cz, pre=-1, cap=4096):
    """One chunk of ``T``'s geometry through libmcgen's CPU path: (density, cells) as ``chunk`` returns them."""
    g = (T.pre_sx, T.pre_sy, T.pre_sz, cx * 16 + T.pre_xoff, T.pre_y0, cz * 16 + T.pre_zoff, T.pre_dx, T.pre_dy, T.pre_dz)
    vol = (16, T.nh, 16, cx * 16, T.nmin, cz * 16, 1, 1, 1)
    cids = [T.cache_ids[k] for k in range(T.ncache_ids)]
    return ref.chunk(density, vol, pre=pre, pre_vol=g, cids=cids, cap=cap)
