"""Builders of McgBiomeWorld descriptions for the plugin tests (libmcgen GPU plugin ABI, include/mcgpu.h).

* ``vanilla(version, dim, seed)``: the oracle's vanilla world in plugin form (or_mcg_world) plus its
  ctypes view;
* ``Prog``: a program assembled node by node in Python (noises borrowed from a vanilla world);
* ``probe_tree(targets)``: a flat R-tree whose leaf j sits exactly on target j with a decoy leaf one unit
  away along each axis, so a search returns j only when the target is reproduced exactly;
* ``with_tree(world, nodes)``: a copy of a world description with another tree.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
for p in (os.path.join(REPO, "oracle"), os.path.join(REPO, "mc-worldgen-seed-lab_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import mcgpu  # noqa: E402  (ctypes structures only; no device call)
import oracle  # noqa: E402

DECOY = 254


class Vanilla:
    def __init__(self, version, dim, seed):
        self.oracle = oracle.Oracle(version, dim)
        self.ref = oracle.McgWorldRef.vanilla(self.oracle, seed)
        self.world = C.cast(C.c_void_p(self.ref.addr), C.POINTER(mcgpu.McgBiomeWorld)).contents
        self.addr = self.ref.addr
        self.seed = seed

    @property
    def prog(self):
        return self.world.prog.contents


def vanilla(version, dim, seed):
    return Vanilla(version, dim, seed)


def _arr(ctype, items):
    n = max(len(items), 1)
    a = (ctype * n)()
    for i, v in enumerate(items):
        a[i] = v
    return a


class Prog:
    """A density-function program written node by node; noises are copied from a vanilla program."""

    def __init__(self, old, src=None):
        self.old = 1 if old else 0
        self.nodes, self.noise, self.layers, self.octs = [], [], [], []
        self.sp, self.splf, self.spc, self.arr, self.thr = [], [], [], [], []
        self.src = src
        self._keep = []

    def node(self, k, **f):
        n = mcgpu.McgNode()
        n.k = k
        n.a = n.b = n.c = -1
        n.ns = n.sp = n.oct = -1
        for key, v in f.items():
            setattr(n, key, v)
        self.nodes.append(n)
        return len(self.nodes) - 1

    def borrow_noise(self, i):
        """Copy noise i of the source program (its layers and octaves) and return its new index."""
        P = self.src
        mn = P.noise[i]
        new = mcgpu.McgNoise()
        new.value_factor = mn.value_factor

        def copy_layers(l0, n):
            first = len(self.layers)
            for q in range(n):
                L = P.layers[l0 + q]
                M = mcgpu.McgLayer(L.a, L.b, L.amp_d, L.amp_f, -1, L.kind, 0)
                if L.oct >= 0:
                    o = P.octs[L.oct]
                    oc = mcgpu.McgOct()
                    oc.xo, oc.yo, oc.zo = o.xo, o.yo, o.zo
                    C.memmove(oc.p, o.p, 256)
                    self.octs.append(oc)
                    M.oct = len(self.octs) - 1
                self.layers.append(M)
            return first

        new.l0, new.n0 = copy_layers(mn.l0, mn.n0), mn.n0
        if self.old:
            new.l1, new.n1 = copy_layers(mn.l1, mn.n1), mn.n1
        self.noise.append(new)
        return len(self.noise) - 1

    def borrow_oct(self, i):
        o = self.src.octs[i]
        oc = mcgpu.McgOct()
        oc.xo, oc.yo, oc.zo = o.xo, o.yo, o.zo
        C.memmove(oc.p, o.p, 256)
        self.octs.append(oc)
        return len(self.octs) - 1

    def spline(self, tree):
        """tree: float constant, or (coord, [(loc, der, subtree), ...]); returns the spline index."""
        idx = len(self.sp)
        self.sp.append(None)
        s = mcgpu.McgSpline()
        if not isinstance(tree, tuple):
            s.is_const, s.value = 1, float(tree)
        else:
            coord, pts = tree
            kids = [self.spline(sub) for _, _, sub in pts]
            s.coord, s.n = coord, len(pts)
            s.loc0 = len(self.splf)
            self.splf += [float(l) for l, _, _ in pts] + [float(d) for _, d, _ in pts]
            s.val0 = len(self.spc)
            self.spc += kids
        self.sp[idx] = s
        return idx

    def list_(self, items):
        a0 = len(self.arr)
        self.arr += list(items)
        return a0

    def build(self, roots):
        P = mcgpu.McgProg()
        P.old = self.old
        parts = [("nodes", mcgpu.McgNode, self.nodes), ("noise", mcgpu.McgNoise, self.noise),
                 ("layers", mcgpu.McgLayer, self.layers), ("octs", mcgpu.McgOct, self.octs),
                 ("sp", mcgpu.McgSpline, self.sp), ("splf", C.c_float, self.splf), ("spc", C.c_int, self.spc),
                 ("arr", C.c_int, self.arr), ("thr", C.c_double, self.thr)]
        counts = {"nodes": "nnodes", "noise": "nnoise", "layers": "nlayers", "octs": "noct", "sp": "nsp",
                  "splf": "nsplf", "spc": "nspc", "arr": "narr", "thr": "nthr"}
        for name, ct, items in parts:
            a = _arr(ct, items)
            self._keep.append(a)
            setattr(P, name, C.cast(a, C.POINTER(ct)))
            setattr(P, counts[name], len(items))
        P.nroots = len(roots)
        for i, r in enumerate(roots):
            P.root[i] = r
        self._keep.append(P)
        return P


def world(prog, like, tree=None, source=mcgpu.MCG_BS_MULTI, keep=None):
    """A McgBiomeWorld around `prog`, taking seed-dependent fields (zoom seed, heights, End ids) from
    the vanilla world description `like`."""
    W = mcgpu.McgBiomeWorld()
    W.source = source
    W.fixed_biome = like.fixed_biome
    for i in range(5):
        W.end_ids[i] = like.end_ids[i]
    W.newf = 0 if prog.old else 1
    W.min_qy, W.qh, W.zoom_seed = like.min_qy, like.qh, like.zoom_seed
    W.prog = C.pointer(prog)
    if tree is not None:
        W.ntree = len(tree)
        W.tree = C.cast(tree, C.POINTER(mcgpu.McgTreeNode))
        W.tree_root = 0
    else:
        W.ntree, W.tree, W.tree_root = like.ntree, like.tree, like.tree_root
    W._keep = (prog, tree, keep)
    return W


def flat_tree(leaves):
    """Root + leaves (lo6, hi6, biome) as direct children, in the given order."""
    n = len(leaves) + 1
    T = (mcgpu.McgTreeNode * n)()
    lo = np.min([l[0] for l in leaves], axis=0)
    hi = np.max([l[1] for l in leaves], axis=0)
    for k in range(6):
        T[0].lo[k], T[0].hi[k] = int(lo[k]), int(hi[k])
    T[0].lo[6] = T[0].hi[6] = 0
    T[0].first, T[0].count, T[0].biome = 1, len(leaves), -1
    for i, (l, h, b) in enumerate(leaves):
        for k in range(6):
            T[i + 1].lo[k], T[i + 1].hi[k] = int(l[k]), int(h[k])
        T[i + 1].lo[6] = T[i + 1].hi[6] = 0
        T[i + 1].first, T[i + 1].count, T[i + 1].biome = -1, 0, int(b)
    return T


def probe_tree(targets):
    """Leaf j exactly at targets[j] (biome j), decoys at +-1 on each axis (biome DECOY)."""
    leaves = []
    for j, t in enumerate(targets):
        t = np.asarray(t, np.int64)
        leaves.append((t, t, j))
    for t in targets:
        for k in range(6):
            for s in (-1, 1):
                d = np.asarray(t, np.int64).copy()
                d[k] += s
                leaves.append((d, d, DECOY))
    return flat_tree(leaves)


def quantize(v):
    """Climate.target's (long)(v * 10000.0f) for one root value (float arithmetic), None when the plugin
    rules reject it."""
    f = np.float32(np.float32(v) * np.float32(10000.0))
    if not (abs(float(f)) < 9.0e18) or not np.isfinite(f):
        return None
    return int(f)


def tree_nodes(W):
    return [W.tree[i] for i in range(W.ntree)]
