"""The R-tree the engine searches, built by the library on the host from a version's parameter list
(mcb_tree; Climate.RTree.create as engine/mc_biomes.h:64-212 and engine/mc_data.h:153-183 build it):
node for node equal to the game's own tree dumps (26.1-26.3, data/rtree_<V>_<dim>.json) and to the
oracle's build, which also covers 26.4-snapshot-2 (upper bounds lowered by one after the build)."""
import json
import os

import numpy as np
import pytest

import oracle as O

mcbiome = pytest.importorskip("mcbiome")

ALL = ["26.1", "26.2", "26.3", "26.4-snapshot-2"]


def _param_biomes(version, dim):
    path = os.path.join(O.DATA_ROOT, "params", version, ("nether" if dim == "the_nether" else "overworld") + ".tsv")
    rows = [l.rstrip("\n").split("\t") for l in open(path) if l.strip() and not l.startswith("#")]
    return [O.biome_id(r[-1].replace("minecraft:", "")) for r in rows]


def _device_rows(rows, biomes):
    """Pre-order rows [is_leaf, lo0, hi0 .. lo6, hi6, param | child count] -> the engine's node layout."""
    out = []

    def walk(i):
        r = rows[i]
        me = len(out)
        out.append(None)
        j = i + 1
        if not r[0]:
            for _ in range(r[15]):
                j = walk(j)
        lo6, hi6 = r[13], r[14]
        x = -hi6 if -hi6 > 0 else (lo6 if lo6 > 0 else 0)          # above = 0 - hi, below = lo - 0
        out[me] = list(r[1:13]) + [x * x, len(out), biomes[r[15]] if r[0] else -1, 0]
        return j

    assert walk(0) == len(rows)
    return np.array(out, np.int64)


@pytest.mark.parametrize("version", ALL)
@pytest.mark.parametrize("dim", ["overworld", "the_nether"])
def test_tree_equals_oracle_build(version, dim):
    got = mcbiome.tree_nodes(version, dim).astype(np.int64)
    want = _device_rows(O.Oracle(version, dim).tree_serialize().tolist(), _param_biomes(version, dim))
    assert got.shape == want.shape and np.array_equal(got, want)


def _json_rows(node, out):
    leaf = "i" in node
    out.append([1 if leaf else 0] + node["b"] + [node["i"] if leaf else len(node["c"])])
    for c in node.get("c", []):
        _json_rows(c, out)


@pytest.mark.parametrize("version", ["26.1", "26.2", "26.3"])
@pytest.mark.parametrize("dim", ["overworld", "the_nether"])
def test_tree_equals_game_dump(version, dim):
    with open(os.path.join(O.DATA_ROOT, "rtree_%s_%s.json" % (version, dim))) as f:
        d = json.load(f)
    rows = []
    _json_rows(d["tree"], rows)
    want = _device_rows(rows, _param_biomes(version, dim))
    got = mcbiome.tree_nodes(version, dim).astype(np.int64)
    assert got.shape[0] == d["leaves"] + d["subtrees"] and np.array_equal(got, want)


def test_snapshot_bounds_are_exclusive():
    """26.4-snapshot-2 differs from 26.3 in two parameter rows and in its upper bounds: every node's hi is
    the 26.3-style build's hi minus one; boxes with lo = hi become lo = hi + 1."""
    t4 = mcbiome.tree_nodes("26.4-snapshot-2", "overworld")
    leaves = t4[t4[:, 14] >= 0]
    assert (leaves[:, 0:12:2] <= leaves[:, 1:12:2] + 1).all()
    assert (leaves[:, 0:12:2] == leaves[:, 1:12:2] + 1).any()
