"""The libmcgen GPU plugin ABI on the host (no compute): every symbol include/mcgpu.h declares is exported,
the ABI version is libmcgen's, the structures have the reference's sizes, and malformed world
descriptions are refused by mcgpu_biome_new with a reason before anything touches the device
(the reference plugin would read out of range on the device instead)."""
import ctypes as C
import os
import re

import pytest

import mcgprog
import mcgpu
import oracle

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "mcgpu.h")


def test_exports_every_declared_symbol():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(mcgpu_[a-z_]+)\s*\(", txt)))
    # the reference plugin's exports (MCGPU_EXPORT in libmcgen/gpu/mcgpu_core.cu, mcgpu_biome.cu, mcgpu_terrain.cu)
    assert names == sorted("""mcgpu_abi mcgpu_biome_free mcgpu_biome_grid mcgpu_biome_new mcgpu_biome_points mcgpu_device_count
        mcgpu_device_info mcgpu_host_alloc mcgpu_host_free mcgpu_last_error mcgpu_mem_info mcgpu_select mcgpu_selected
        mcgpu_terrain_batch mcgpu_terrain_cell_cand mcgpu_terrain_cell_floats mcgpu_terrain_cell_info mcgpu_terrain_free
        mcgpu_terrain_max_chunks mcgpu_terrain_new""".split())
    L = mcgpu.lib()
    assert not [n for n in names if not hasattr(L, n)]
    assert L.mcgpu_abi() == mcgpu.MCGPU_ABI == 2


def test_struct_sizes_match_reference_layout():
    # libmcgen/gpu/mcgen_gpu_abi.h on LP64: field order and types fix these sizes
    assert C.sizeof(mcgpu.McgOct) == 280
    assert C.sizeof(mcgpu.McgLayer) == 40
    assert C.sizeof(mcgpu.McgNoise) == 24
    assert C.sizeof(mcgpu.McgSpline) == 24
    assert C.sizeof(mcgpu.McgNode) == 112
    assert C.sizeof(mcgpu.McgTreeNode) == 72
    assert C.sizeof(mcgpu.McgBiomeWorld) == 80
    assert C.sizeof(mcgpu.McgTerrainWorld) == 336 and mcgpu.McgTerrainWorld.ore_lo.offset == 144
    assert mcgpu.K["FTS"] == 56 and mcgpu.O["PASS"] == 28


def _refused(W, text):
    L = mcgpu.lib()
    err = C.create_string_buffer(512)
    h = L.mcgpu_biome_new(C.c_void_p(C.addressof(W)), err, len(err))
    assert not h
    assert text in err.value.decode(), err.value.decode()


@pytest.fixture(scope="module")
def van():
    return mcgprog.vanilla("26.3", "overworld", 1)


def _const_prog(P, roots_from=None):
    c = P.node(mcgpu.K["CONST"], f0=0.25)
    return [c] * 6


def test_refuses_unsupported_kind(van):
    P = mcgprog.Prog(False, van.prog)
    c = P.node(mcgpu.K["CONST"], f0=2.0)
    p = P.node(mcgpu.K["POW"], a=c, b=c)
    _refused(mcgprog.world(P.build([p] * 6), van.world), "not supported")


def test_refuses_forward_reference(van):
    P = mcgprog.Prog(False, van.prog)
    a = P.node(mcgpu.K["ABS"], a=1)
    P.node(mcgpu.K["CONST"], f0=1.0)
    _refused(mcgprog.world(P.build([a] * 6), van.world), "must be an earlier node")


def test_refuses_missing_field(van):
    P = mcgprog.Prog(False, van.prog)
    c = P.node(mcgpu.K["CONST"], f0=1.0)
    _refused(mcgprog.world(P.build([c, c, c, -1, c, c]), van.world), "no climate field 3")


def test_refuses_branch_mismatch(van):
    P = mcgprog.Prog(True, van.prog)
    c = P.node(mcgpu.O["CONST"], d0=1.0)
    W = mcgprog.world(P.build([c] * 6), van.world)
    W.newf = 1
    _refused(W, "does not match")


def test_refuses_deep_spline(van):
    P = mcgprog.Prog(False, van.prog)
    x = P.node(mcgpu.K["CONST"], f0=0.5)
    t = 1.0
    for _ in range(17):
        t = (0, [(0.0, 0.0, t), (1.0, 0.0, 2.0)])
    sp = P.spline(t)
    s = P.node(mcgpu.K["SPLINE"], sp=sp, arr0=P.list_([x]), narr=1)
    _refused(mcgprog.world(P.build([s] * 6), van.world), "deeper than 16")


def test_refuses_isel_thresholds(van):
    P = mcgprog.Prog(False, van.prog)
    x = P.node(mcgpu.K["CONST"], f0=0.5)
    P.thr += [0.0, 1.0]
    s = P.node(mcgpu.K["ISEL"], a=x, arr0=P.list_([x, x]), narr=2, thr0=0, nthr=2)
    _refused(mcgprog.world(P.build([s] * 6), van.world), "more thresholds than choices")


def test_refuses_noise_out_of_range(van):
    P = mcgprog.Prog(False, van.prog)
    n = P.node(mcgpu.K["NOISE"], ns=3, d0=1.0, d1=1.0)
    _refused(mcgprog.world(P.build([n] * 6), van.world), "noise index 3 out of range")


def test_refuses_tree_cycle(van):
    P = mcgprog.Prog(False, van.prog)
    c = P.node(mcgpu.K["CONST"], f0=0.5)
    T = mcgprog.flat_tree([([0] * 6, [0] * 6, 1), ([1] * 6, [1] * 6, 2)])
    T[1].first, T[1].count = 0, 1            # a leaf turned into a node whose child is the root
    _refused(mcgprog.world(P.build([c] * 6), van.world, tree=T), "reached twice")


def test_refuses_zero_gradient_period(van):
    P = mcgprog.Prog(False, van.prog)
    g = P.node(mcgpu.K["GRAD_REPEAT"], i0=1, i1=0, i2=0, f0=0.0, f1=1.0)
    _refused(mcgprog.world(P.build([g] * 6), van.world), "gradient period 0")


# ---- recognition of the vanilla Overworld climate router (host only, no device call) ----
def _vanilla_world(version, preset, seed=11):
    o = oracle.Oracle(version, "overworld", preset)
    ref = oracle.McgWorldRef.vanilla(o, seed)
    W = C.cast(C.c_void_p(ref.addr), C.POINTER(mcgpu.McgBiomeWorld)).contents
    return ref, W


def _nodes_of(W, kind):
    P = W.prog.contents
    return [i for i in range(P.nnodes) if P.nodes[i].k == kind]


@pytest.mark.parametrize("version", ["26.1", "26.2", "26.3"])
@pytest.mark.parametrize("preset", ["overworld", "large_biomes", "amplified"])
def test_router_recognises_vanilla_overworld(version, preset):
    ref, W = _vanilla_world(version, preset)
    assert mcgpu.router_world(ref.addr)


@pytest.mark.parametrize("dim", ["nether", "end"])
def test_router_not_for_other_dimensions(dim):
    V = mcgprog.vanilla("26.3", dim, 3)
    assert not mcgpu.router_world(V.addr)


def _perturb_new(W, what):
    """One change to a 26.3 vanilla program that the engine's column code does not compute."""
    P = W.prog.contents
    K = mcgpu.K
    if what == "shift_scale":
        i = [i for i in _nodes_of(W, K["CMUL"]) if P.nodes[i].f0 == 4.0][0]
        P.nodes[i].f0 = 4.5
    elif what == "alpha":
        P.nodes[_nodes_of(W, K["CTX_ALPHA"])[0]].f0 = 0.5
    elif what == "second_freq":
        L = P.layers[P.noise[3].l0 + 3]                 # continentalness, level 1, second stack
        L.a = L.a * (1.0 + 2.0 ** -40)
    elif what == "smeared":
        P.layers[P.noise[1].l0].kind = 1
    elif what == "y_scale":
        P.nodes[_nodes_of(W, K["NOISE_XZ"])[2]].d1 = 0.5
    elif what == "slice":
        n = P.nodes[_nodes_of(W, K["NOISE_XZ"])[0]]
        n.cm, n.cx = 1, 0
    elif what == "levels":
        P.noise[4].n0 -= 2                               # erosion with one level fewer
    elif what == "depth_axis":
        P.nodes[_nodes_of(W, K["GRAD_CLAMP"])[0]].i0 = 0
    elif what == "ridge_root":
        P.root[5] = P.root[2]                            # ridges = continentalness
    else:
        raise ValueError(what)


@pytest.mark.parametrize("what", ["shift_scale", "alpha", "second_freq", "smeared", "y_scale", "slice", "levels",
                                  "depth_axis", "ridge_root"])
def test_router_refuses_other_graphs(what):
    ref, W = _vanilla_world("26.3", "overworld")
    assert mcgpu.router_world(ref.addr)
    _perturb_new(W, what)
    assert not mcgpu.router_world(ref.addr)


def test_router_refuses_other_old_graphs():
    ref, W = _vanilla_world("26.1", "overworld")
    P = W.prog.contents
    O = mcgpu.O
    i = _nodes_of(W, O["SHIFTED_NOISE"])[1]
    P.nodes[i].b = _nodes_of(W, O["SHIFT_A"])[0]         # y shifted by shift_x instead of 0
    assert not mcgpu.router_world(ref.addr)
    ref, W = _vanilla_world("26.2", "overworld")
    P = W.prog.contents
    P.layers[P.noise[2].l1 + 1].b *= 2.0                 # second PerlinNoise's value factor differs
    assert not mcgpu.router_world(ref.addr)


def test_router_maps_spline_coordinates():
    """The spline's coordinate list permuted (and every spline node's coordinate renumbered to match) is
    the same function; the recognition follows the list instead of assuming its order."""
    ref, W = _vanilla_world("26.3", "overworld")
    P = W.prog.contents
    s = P.nodes[_nodes_of(W, mcgpu.K["SPLINE"])[0]]
    perm = [2, 0, 3, 1]                                  # new position of each old coordinate
    old = [P.arr[s.arr0 + j] for j in range(4)]
    for j in range(4):
        P.arr[s.arr0 + perm[j]] = old[j]
    for i in range(P.nsp):
        if not P.sp[i].is_const:
            P.sp[i].coord = perm[P.sp[i].coord]
    assert mcgpu.router_world(ref.addr)
