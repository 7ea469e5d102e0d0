"""Parity of the libmcgen GPU plugin entry points (mcgpu_*, libmcbiome_hip.so) with oracle/mcgplugin.c.

* vanilla worlds (26.3 / 26.1, Overworld / Nether / End, built by or_mcg_world): points anywhere, y
  beyond the world's height (clamped), grids through the cell-lattice path (step 1, 2), the per-point
  path (step 16) and the End chunk table; outputs and unresolved counts equal;
* every node kind of both program branches, through probe trees: leaf j sits exactly on the oracle's
  quantized target of probe point j (values scaled by an exact power of two first) and decoy leaves one
  unit away on every axis, so the device returns j only when it reproduces every climate value exactly;
* worlds with overlapping leaves (ties everywhere, chains long enough to be refused), NaN climate
  values, the fixed source.
"""
import ctypes as C

import numpy as np
import pytest

import mcgprog
import oracle

pytestmark = pytest.mark.gpu
mcgpu = pytest.importorskip("mcgpu")
K, O = mcgpu.K, mcgpu.O


@pytest.fixture(scope="module", autouse=True)
def _device():
    if mcgpu.lib().mcgpu_device_count() < 1:
        pytest.fail("no HIP device visible to libmcbiome_hip.so")


def test_device_entry_points():
    L = mcgpu.lib()
    assert L.mcgpu_select(0) == 0 and L.mcgpu_selected() == 0
    name = C.create_string_buffer(128)
    ma, mi, mem = C.c_int(), C.c_int(), C.c_size_t()
    assert L.mcgpu_device_info(0, name, 128, C.byref(ma), C.byref(mi), C.byref(mem)) == 0
    print("device", name.value, ma.value, mi.value, mem.value, flush=True)
    assert mem.value > 1024
    f, t = C.c_size_t(), C.c_size_t()
    assert L.mcgpu_mem_info(C.byref(f), C.byref(t)) == 0 and 0 < f.value <= t.value


def _points(rng, n, span, ylo, yhi):
    xyz = np.empty((n, 3), np.int32)
    xyz[:, 0] = rng.integers(-span, span, n)
    xyz[:, 2] = rng.integers(-span, span, n)
    xyz[:, 1] = rng.integers(ylo, yhi, n)
    return xyz


HEIGHT = {"overworld": (-64, 320), "nether": (0, 256), "end": (0, 256)}
CASES = [(v, d) for v in ("26.3", "26.1") for d in ("overworld", "nether", "end")]


@pytest.mark.parametrize("version,dim", CASES)
def test_vanilla_points(version, dim):
    V = mcgprog.vanilla(version, dim, -7 if dim != "end" else 12345)
    rng = np.random.default_rng(hash((version, dim)) & 0xFFFF)
    lo, hi = HEIGHT[dim]
    xyz = np.concatenate([_points(rng, 3000, 3_000_000, lo - 40, hi + 40), _points(rng, 1000, 3000, lo, hi)])
    print("points", version, dim, len(xyz), flush=True)
    W = mcgpu.BiomeWorld(V.world)
    got, bad = W.points(xyz)
    want, wbad = V.ref.points(xyz)
    assert bad == wbad
    assert np.array_equal(got, want), int((got != want).sum())


@pytest.mark.parametrize("version,dim", CASES)
@pytest.mark.parametrize("step,n", [(1, 150), (2, 90), (16, 60)])
def test_vanilla_grid(version, dim, step, n):
    V = mcgprog.vanilla(version, dim, 3)
    x0, z0 = (-1203, 2051) if dim != "end" else (1800, -2207)
    y = {"overworld": 70, "nether": 40, "end": 60}[dim]
    print("grid", version, dim, step, n, flush=True)
    W = mcgpu.BiomeWorld(V.world)
    got, bad = W.grid(x0, z0, n + 7, n, step, y)
    want, wbad = V.ref.grid(x0, z0, n + 7, n, step, y)
    assert bad == wbad
    assert np.array_equal(got, want), int((got != want).sum())


def test_large_grid_sampled():
    V = mcgprog.vanilla("26.3", "overworld", 77)
    W = mcgpu.BiomeWorld(V.world)
    n = 1024
    got, bad = W.grid(-40_000, 90_000, n, n, 1, 100)
    assert bad == 0
    rng = np.random.default_rng(9)
    iz, ix = rng.integers(0, n, 3000), rng.integers(0, n, 3000)
    xyz = np.stack([-40_000 + ix, np.full(3000, 100), 90_000 + iz], 1).astype(np.int32)
    want, _ = V.ref.points(xyz)
    assert np.array_equal(got[iz, ix], want)
    # a second call on the same world (buffers reused) gives the same answer
    again, _ = W.grid(-40_000, 90_000, n, n, 1, 100)
    assert np.array_equal(got, again)


@pytest.mark.parametrize("version,dim,step", [("26.3", "overworld", 1), ("26.3", "overworld", 4), ("26.1", "overworld", 2),
                                              ("26.3", "end", 1), ("26.3", "nether", 16)])
def test_large_grid_equals_row_chunks(version, dim, step):
    """A grid of more than 1 MiB (one lattice over all its quart rows, or one per-point launch) equals
    the same rows requested as grids of 128 rows, whose lattices start at other quart rows; the row
    count is not a multiple of 128 and a chunk edge falls inside a quart row."""
    V = mcgprog.vanilla(version, dim, 5)
    W = mcgpu.BiomeWorld(V.world)
    nx, nz = 1100, 1031
    x0, z0, y = (-3001, 70_003, 64) if dim != "end" else (2203, -1905, 60)
    print("large grid", version, dim, step, flush=True)
    got, bad = W.grid(x0, z0, nx, nz, step, y)
    parts, pbad = [], 0
    for r in range(0, nz, 128):
        g, b = W.grid(x0, z0 + r * step, nx, min(128, nz - r), step, y)
        parts.append(g)
        pbad += b
    want = np.concatenate(parts, 0)
    assert bad == pbad
    assert np.array_equal(got, want), int((got != want).sum())


# ---------------------------------------------------------------------------------------------------
# node kinds through probe trees
# ---------------------------------------------------------------------------------------------------
def kitchen_new(src):
    P = mcgprog.Prog(False, src)
    t, c, r = P.borrow_noise(1), P.borrow_noise(3), P.borrow_noise(5)
    sm = P.borrow_noise(2)
    for i in range(P.noise[sm].n0):
        L = P.layers[P.noise[sm].l0 + i]
        L.kind, L.b = 1, 0.37 + 0.1 * i
    n = {}
    n["x0"] = P.node(K["NOISE"], ns=t, d0=0.25, d1=0.25)
    n["x1"] = P.node(K["NOISE"], ns=c, d0=1.0 / 64, d1=0.0)
    n["smear"] = P.node(K["NOISE"], ns=sm, d0=0.5, d1=0.7)
    n["xz"] = P.node(K["NOISE_XZ"], ns=t, d0=0.25, d1=0.1, a=n["x1"], c=n["x0"])
    n["xyz"] = P.node(K["NOISE_XYZ"], ns=c, d0=0.3, d1=0.2, a=n["x0"], b=n["x1"], c=n["xz"])
    n["shiftb"] = P.node(K["SHIFT_B"], ns=r, d0=0.25)
    n["gy"] = P.node(K["GRAD_CLAMP"], i0=1, i1=-64, i2=-64, i3=320, f0=1.5, f1=-3.0 / 384)
    n["gx"] = P.node(K["GRAD_REPEAT"], i0=0, i1=5, i2=37, f0=-1.0, f1=0.05)
    n["gz"] = P.node(K["GRAD_MIRROR"], i0=2, i1=-3, i2=-29, f0=0.5, f1=0.03)
    for m in range(4):
        n["dist%d" % m] = P.node(K["DIST"], i0=100, i1=64, i2=-200, i3=m)
    sl = P.node(K["NOISE"], ns=t, d0=0.25, d1=0.25, cm=2, cy=40)
    n["slice_y"] = P.node(K["SLICE_Y"], a=sl)
    sx = P.node(K["DIST"], i0=3, i1=5, i2=7, i3=0, cm=5, cx=11, cz=-13)
    n["slice_xz"] = P.node(K["SLICE_XZ"], a=sx)
    x0, x1, xz = n["x0"], n["x1"], n["xz"]
    for name, k in [("abs", "ABS"), ("square", "SQUARE"), ("cube", "CUBE"), ("neg", "NEG"), ("sign", "SIGN"),
                    ("recip", "RECIP"), ("cache", "CACHE"), ("blend", "BLEND_DENSITY")]:
        n[name] = P.node(K[k], a=x0 if name != "recip" else x1)
    n["sqrt"] = P.node(K["SQRT"], a=n["abs"])
    n["sqrt_neg"] = P.node(K["SQRT"], a=x0)                # NaN where x0 < 0: unresolved
    n["leaky"] = P.node(K["LEAKY"], a=x0, f0=0.3)
    n["squeeze"] = P.node(K["SQUEEZE"], a=P.node(K["CMUL"], a=x0, f0=3.0))
    n["add"] = P.node(K["ADD"], a=x0, b=x1)
    n["cadd"] = P.node(K["CADD"], a=x0, f0=0.7)
    n["csub"] = P.node(K["CSUB"], a=x1, f0=0.3)
    n["sub"] = P.node(K["SUB"], a=x0, b=x1)
    n["mul"] = P.node(K["MUL"], a=x0, b=xz)
    n["cmul"] = P.node(K["CMUL"], a=x1, f0=-2.5)
    n["div"] = P.node(K["DIV"], a=x0, b=x1)
    n["cdiv"] = P.node(K["CDIV"], a=x0, f0=1.5)
    n["min"] = P.node(K["MIN"], a=x0, b=x1, f0=-0.2)
    n["cmin"] = P.node(K["CMIN"], a=x0, f0=0.1)
    n["max"] = P.node(K["MAX"], a=x0, b=x1, f0=0.3)
    n["cmax"] = P.node(K["CMAX"], a=x1, f0=-0.1)
    alpha = P.node(K["CLAMP"], a=P.node(K["CMUL"], a=x1, f0=3.0), f0=0.0, f1=1.0)
    n["lerp"] = P.node(K["LERP"], a=alpha, b=x0, c=xz)
    n["lerp_cf"] = P.node(K["LERP_CF"], a=alpha, c=x0, f0=0.25)
    n["lerp_cs"] = P.node(K["LERP_CS"], a=alpha, b=x0, f0=-0.5)
    n["clamp"] = P.node(K["CLAMP"], a=x0, f0=-0.3, f1=0.4)
    n["range_c"] = P.node(K["RANGE_C"], a=x0, f0=-0.2, f1=0.2, f2=1.5, f3=-2.5)
    n["range"] = P.node(K["RANGE"], a=x0, f0=-0.1, f1=0.3, b=x1, c=xz)
    n["isel1"] = P.node(K["ISEL1"], a=x1, f0=0.05, b=x0, c=xz)
    P.thr += [-0.3, 0.0, 0.2]
    n["isel"] = P.node(K["ISEL"], a=x1, thr0=0, nthr=3, arr0=P.list_([x0, xz, n["xyz"], n["shiftb"]]), narr=4)
    for k in ("CTX_ALPHA", "CTX_OFFSET", "CTX_BEARD"):
        n[k.lower()] = P.node(K[k], f0=0.625)
    sp = P.spline((0, [(-0.5, 0.1, 0.2), (-0.1, 0.0, (1, [(-0.3, 0.5, -0.1), (0.2, -1.0, 0.4)])), (0.3, 2.0, 0.9)]))
    n["spline"] = P.node(K["SPLINE"], sp=sp, arr0=P.list_([x0, x1]), narr=2)
    n["end"] = P.node(K["END"], oct=P.borrow_oct(0))
    return P, n


def kitchen_old(src):
    P = mcgprog.Prog(True, src)
    t, c, o = P.borrow_noise(1), P.borrow_noise(3), P.borrow_noise(0)
    n = {}
    n["x0"] = P.node(O["NOISE"], ns=t, d0=0.25, d1=0.25)
    n["x1"] = P.node(O["NOISE"], ns=c, d0=1.0 / 64, d1=0.0)
    x0, x1 = n["x0"], n["x1"]
    n["shifted"] = P.node(O["SHIFTED_NOISE"], ns=t, d0=0.25, d1=0.5, a=x1, b=x0, c=x1)
    n["shift_a"] = P.node(O["SHIFT_A"], ns=o, d0=0.25)
    n["shift_b"] = P.node(O["SHIFT_B"], ns=o, d0=0.25)
    n["shift"] = P.node(O["SHIFT"], ns=o, d0=0.25, d1=0.25)
    n["ygrad"] = P.node(O["Y_GRADIENT"], i0=-64, i1=320, d0=1.5, d1=-1.5)
    n["weird0"] = P.node(O["WEIRD_SCALED"], ns=c, a=x0, i0=0)
    n["weird1"] = P.node(O["WEIRD_SCALED"], ns=c, a=x1, i0=1)
    n["alpha"] = P.node(O["BLEND_ALPHA"])
    n["zero"] = P.node(O["ZERO"])
    for name in ("ABS", "SQUARE", "CUBE", "HALF_NEG", "QUARTER_NEG", "SQUEEZE", "PASS"):
        n[name.lower()] = P.node(O[name], a=x0)
    n["recip"] = P.node(O["RECIP"], a=x1)
    n["clamp"] = P.node(O["CLAMP"], a=x0, d0=-0.3, d1=0.4)
    n["add"] = P.node(O["ADD"], a=x0, b=x1)
    n["mul"] = P.node(O["MUL"], a=x0, b=x1)
    n["min"] = P.node(O["MIN"], a=x0, b=x1, d1=-0.2)
    n["max"] = P.node(O["MAX"], a=x0, b=x1, d1=0.3)
    n["cadd"] = P.node(O["CADD"], a=x0, d0=0.7)
    n["cmul"] = P.node(O["CMUL"], a=x1, d0=-2.5)
    n["range"] = P.node(O["RANGE"], a=x0, d0=-0.1, d1=0.3, b=x1, c=n["shifted"])
    P.thr += [-0.3, 0.0, 0.2]
    n["isel"] = P.node(O["ISEL"], a=x1, thr0=0, nthr=3, arr0=P.list_([x0, x1, n["shifted"], n["shift"]]), narr=4)
    sp = P.spline((0, [(-0.5, 0.1, 0.2), (-0.1, 0.0, (1, [(-0.3, 0.5, -0.1), (0.2, -1.0, 0.4)])), (0.3, 2.0, 0.9)]))
    n["spline"] = P.node(O["SPLINE"], sp=sp, arr0=P.list_([x0, x1]), narr=2)
    n["end"] = P.node(O["END_ISLANDS"], oct=P.borrow_oct(0))
    return P, n


def deep_spline_new(src):
    P = mcgprog.Prog(False, src)
    x0 = P.node(K["NOISE"], ns=P.borrow_noise(1), d0=0.25, d1=0.25)
    tree = 0.5
    for d in range(9):
        tree = (0, [(-0.4 + 0.05 * d, 0.3, tree), (0.1 * d - 0.2, -0.7, 1.0 - 0.1 * d), (0.6, 0.0, tree)])
    sp = P.spline(tree)
    return P, {"deep": P.node(K["SPLINE"], sp=sp, arr0=P.list_([x0]), narr=1)}


def _scale(P, node, scale, old):
    return P.node(O["CMUL"], a=node, d0=float(scale)) if old else P.node(K["CMUL"], a=node, f0=float(scale))


def probe(V, P, nodes, old, xyz):
    """Checks every node of `nodes` (6 at a time as the world's climate fields) on the points xyz."""
    names = list(nodes)
    hs = int(V.world.zoom_seed)
    for g in range(0, len(names), 6):
        group = names[g:g + 6]
        group += [group[-1]] * (6 - len(group))
        roots = [nodes[k] for k in group]
        # scale each field by a power of two (exact) so its quantized value resolves single ulps
        plain = mcgprog.world(P.build(roots), V.world)
        ref0 = oracle.McgWorldRef(C.addressof(plain), keep=plain)
        lo_q, hi_q = V.world.min_qy, V.world.min_qy + V.world.qh - 1
        cells = [(q[0], min(max(q[1], lo_q), hi_q), q[2]) for q in (oracle.zoom(hs, *p) for p in xyz)]
        vals = np.array([ref0.climate(4 * c[0], 4 * c[1], 4 * c[2]) for c in cells])
        finite = np.where(np.isfinite(vals), np.abs(vals), 0.0)
        peak = np.maximum(finite.max(axis=0), 1e-3)
        scale = [2.0 ** min(13, int(np.floor(np.log2(1.5e9 / (1e4 * p))))) for p in peak]
        roots = [_scale(P, r, s, old) for r, s in zip(roots, scale)]
        prog = P.build(roots)
        plain = mcgprog.world(prog, V.world)
        ref0 = oracle.McgWorldRef(C.addressof(plain), keep=plain)
        targets, keep, taken = {}, [], set()
        for i, c in enumerate(cells):
            qy = c[1]
            v = ref0.climate(4 * c[0], 4 * qy, 4 * c[2])
            t = [mcgprog.quantize(x) for x in v]
            # a probe needs its own target: not a value another probe (or its decoys) already holds
            near = [] if any(x is None for x in t) else [tuple(t)] + [
                tuple(t[:k] + [t[k] + s] + t[k + 1:]) for k in range(6) for s in (-1, 1)]
            if not near or len(targets) >= 200 or (c[0], qy, c[2]) in targets or any(q in taken for q in near):
                keep.append((i, None))
                continue
            taken.update(near)
            targets[(c[0], qy, c[2])] = len(targets)
            keep.append((i, targets[(c[0], qy, c[2])]))
        tlist = [None] * len(targets)
        for (cx, cy, cz), j in targets.items():
            tlist[j] = [mcgprog.quantize(x) for x in ref0.climate(4 * cx, 4 * cy, 4 * cz)]
        W = mcgprog.world(prog, V.world, tree=mcgprog.probe_tree(tlist))
        ref = oracle.McgWorldRef(C.addressof(W), keep=W)
        sel = np.array([i for i, j in keep if j is not None])
        exp = np.array([j for i, j in keep if j is not None], np.uint8)
        want, _ = ref.points(xyz[sel])
        assert np.array_equal(want, exp), "probe construction"      # the oracle finds its own targets
        print("probe", group, len(sel), flush=True)
        got, _ = mcgpu.BiomeWorld(W).points(xyz[sel])
        assert np.array_equal(got, exp), (group, int((got != exp).sum()))
        # positions the plugin rules reject (NaN / huge values) come back unresolved on both sides
        rej = np.array([i for i, j in keep if j is None], np.int64)
        if len(rej):
            g2, b2 = mcgpu.BiomeWorld(W).points(xyz[rej])
            w2, wb2 = ref.points(xyz[rej])
            assert np.array_equal(g2, w2) and b2 == wb2


@pytest.mark.parametrize("builder,version,old", [(kitchen_new, "26.3", False), (kitchen_old, "26.1", True),
                                                 (deep_spline_new, "26.3", False)])
def test_node_kinds_exact(builder, version, old):
    V = mcgprog.vanilla(version, "overworld", 31)
    P, nodes = builder(V.prog)
    rng = np.random.default_rng(17)
    xyz = np.concatenate([_points(rng, 250, 40_000, -64, 320), _points(rng, 50, 5_000_000, -64, 320)])
    probe(V, P, nodes, old, xyz)


@pytest.mark.parametrize("builder,version,old", [(kitchen_new, "26.3", False), (kitchen_old, "26.1", True)])
def test_node_kinds_lattice_grid(builder, version, old):
    """Step-1 grids evaluate both quart layers of a lattice column in one lane, the second skipping the
    nodes whose value cannot depend on y.  Every node kind, as a climate field over the vanilla tree,
    must give the oracle's grid, including the y-dependent ones (y gradients, distances, noises with a y
    scale) and everything computed from them.  Fields are scaled by powers of two as in the probes, so no
    squared parameter distance leaves the 64-bit range: beyond it the plugin ABI's unsigned sums
    (mcgpu_dev.cuh:464-473) and libmcgen's signed ones, which the oracle follows, part ways."""
    V = mcgprog.vanilla(version, "overworld", 31)
    P, nodes = builder(V.prog)
    names = list(nodes)
    x0, z0, nx, nz, y = -1203, 2051, 67, 60, 70
    cells = [(qx, qy, qz) for qx in range((x0 - 2) >> 2, ((x0 + nx - 3) >> 2) + 2, 5)
             for qz in range((z0 - 2) >> 2, ((z0 + nz - 3) >> 2) + 2, 5) for qy in ((y - 2) >> 2, ((y - 2) >> 2) + 1)]
    for g in range(0, len(names), 6):
        group = names[g:g + 6]
        group += [group[-1]] * (6 - len(group))
        roots = [nodes[k] for k in group]
        plain = mcgprog.world(P.build(roots), V.world)
        ref0 = oracle.McgWorldRef(C.addressof(plain), keep=plain)
        vals = np.array([ref0.climate(4 * c[0], 4 * c[1], 4 * c[2]) for c in cells])
        peak = np.maximum(np.where(np.isfinite(vals), np.abs(vals), 0.0).max(axis=0), 1e-3)
        scale = [2.0 ** min(13, int(np.floor(np.log2(1.5e9 / (1e4 * p))))) for p in peak]
        W = mcgprog.world(P.build([_scale(P, r, s, old) for r, s in zip(roots, scale)]), V.world)
        ref = oracle.McgWorldRef(C.addressof(W), keep=W)
        print("lattice grid", group, flush=True)
        got, bad = mcgpu.BiomeWorld(W).grid(x0, z0, nx, nz, 1, y)
        want, wbad = ref.grid(x0, z0, nx, nz, 1, y)
        assert bad == wbad, group
        assert np.array_equal(got, want), (group, int((got != want).sum()))


# ---------------------------------------------------------------------------------------------------
# ties, NaN, fixed source
# ---------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("shift", [1500, 6000, "snapped"])
def test_overlapping_leaves_chain(shift):
    """Walk path (copies add breakpoints) and cell-table path (snapped copies reuse them)."""
    from test_oracle_plugin import overlap_world, overlap_world_snapped
    V, W = overlap_world_snapped("26.1", 8) if shift == "snapped" else overlap_world("26.3", 8, shift)
    shift = 7 if shift == "snapped" else shift
    ref = oracle.McgWorldRef(C.addressof(W), keep=W)
    rng = np.random.default_rng(shift)
    xyz = _points(rng, 800, 200_000, -64, 320)
    want, wbad = ref.points(xyz)
    got, bad = mcgpu.BiomeWorld(W).points(xyz)
    print("ties", shift, "unresolved", wbad, flush=True)
    assert bad == wbad
    assert np.array_equal(got, want), int((got != want).sum())
    g, gb = mcgpu.BiomeWorld(W).grid(1000, -3000, 48, 40, 1, 30)
    w, wb = ref.grid(1000, -3000, 48, 40, 1, 30)
    assert gb == wb and np.array_equal(g, w)


def test_fixed_source():
    V = mcgprog.vanilla("26.3", "overworld", 1)
    W = mcgprog.world(V.prog, V.world, source=mcgpu.MCG_BS_FIXED)
    W.fixed_biome = 41
    g, bad = mcgpu.BiomeWorld(W).grid(-50, 70, 33, 17, 5, 12)
    assert bad == 0 and (g == 41).all()


# ---------------------------------------------------------------------------------------------------
# the engine's exporter (mcb_plugin_world) writes the oracle's plugin-form world, array for array
# ---------------------------------------------------------------------------------------------------
def _fields(s):
    return [(n, getattr(s, n)) for n, _ in s._fields_]


def _eq_struct(a, b, skip=()):
    for (n, x), (_, y) in zip(_fields(a), _fields(b)):
        if n in skip:
            continue
        if hasattr(x, "_length_"):
            assert list(x) == list(y), n
        else:
            assert x == y or (x != x and y != y), n


@pytest.mark.parametrize("version,dim", CASES)
def test_export_equals_oracle_world(version, dim):
    import mcbiome
    for seed in (0, -4242424242):
        V = mcgprog.vanilla(version, dim, seed)
        e = mcbiome.Engine(version, dim)
        pw = e.plugin_world(seed)
        W = C.cast(C.c_void_p(pw.addr), C.POINTER(mcgpu.McgBiomeWorld)).contents
        R = V.world
        for n in ("source", "newf", "min_qy", "qh", "zoom_seed", "ntree", "tree_root"):
            assert getattr(W, n) == getattr(R, n), n
        assert list(W.end_ids) == list(R.end_ids)
        P, Q = W.prog.contents, R.prog.contents
        for n in ("old", "nnodes", "nnoise", "nlayers", "noct", "nsp", "nsplf", "nspc", "narr", "nroots"):
            assert getattr(P, n) == getattr(Q, n), n
        assert list(P.root)[:6] == list(Q.root)[:6]
        for arr, cnt in (("nodes", "nnodes"), ("noise", "nnoise"), ("layers", "nlayers"), ("sp", "nsp")):
            for i in range(getattr(P, cnt)):
                _eq_struct(getattr(P, arr)[i], getattr(Q, arr)[i])
        for i in range(P.noct):
            a, b = P.octs[i], Q.octs[i]
            assert (a.xo, a.yo, a.zo) == (b.xo, b.yo, b.zo) and bytes(a.p) == bytes(b.p), i
        assert [P.splf[i] for i in range(P.nsplf)] == [Q.splf[i] for i in range(Q.nsplf)]
        assert [P.spc[i] for i in range(P.nspc)] == [Q.spc[i] for i in range(Q.nspc)]
        assert [P.arr[i] for i in range(P.narr)] == [Q.arr[i] for i in range(Q.narr)]
        for i in range(W.ntree):
            a, b = W.tree[i], R.tree[i]
            _eq_struct(a, b, skip=("lo", "hi"))
            assert list(a.lo)[:6] == list(b.lo)[:6] and list(a.hi)[:6] == list(b.hi)[:6], i
            # offset dimension: the same squared distance to a target offset of 0
            d = max(-b.hi[6], b.lo[6], 0) if -b.hi[6] <= 0 else -b.hi[6]
            assert a.lo[6] == a.hi[6] == d, i
        # and the plugin gives the same biomes from either description
        rng = np.random.default_rng(seed & 0xFFFF)
        lo, hi = HEIGHT[dim]
        xyz = _points(rng, 500, 1_000_000, lo, hi)
        g1, _ = mcgpu.BiomeWorld(pw.addr).points(xyz)
        g2, _ = mcgpu.BiomeWorld(V.world).points(xyz)
        assert np.array_equal(g1, g2)
