"""The climate cell table (the per-sample replacement of the R-tree search) through mcb_targets: the
device result equals the oracle's tree search on targets on, beside and between every breakpoint,
inside every kind of cell, in the outer cells past +-32767, beyond the table's +-2^17 range (tree-walk path) and at the API limit."""
import numpy as np
import pytest

import golden_io as G
import oracle as O
from test_oracle_targets import adversarial_targets, breakpoints

pytestmark = pytest.mark.gpu
mcbiome = pytest.importorskip("mcbiome")


@pytest.fixture(scope="module", autouse=True)
def _device():
    if mcbiome.device_count() < 1:
        pytest.fail("no HIP device visible to libmcbiome_hip.so")


@pytest.mark.parametrize("version", G.VERSIONS)
def test_cell_table_targets_match_oracle(version):
    eng = mcbiome.Engine(version, "overworld")
    cells, list_len, ms = eng.cells()
    print("cells", version, cells, list_len, "build ms", round(ms, 1), flush=True)
    assert cells > 0 and list_len >= cells
    orc = O.Oracle(version, "overworld")
    rng = np.random.default_rng(40 + G.VERSIONS.index(version))
    tg = adversarial_targets(orc, rng, 200000)
    got = eng.targets(tg)
    want = orc.targets(tg)
    assert np.array_equal(got, want), int((got != want).sum())


@pytest.mark.parametrize("version", G.VERSIONS)
def test_cell_corners_match_oracle(version):
    """Every combination of breakpoint-adjacent values in two dimensions, the rest at cell centres."""
    eng = mcbiome.Engine(version, "overworld")
    orc = O.Oracle(version, "overworld")
    bps = breakpoints(orc)
    rng = np.random.default_rng(77)
    rows = []
    for a in range(6):
        for b in range(a + 1, 6):
            va = np.unique(np.concatenate([bps[a], bps[a] - 1]))
            vb = np.unique(np.concatenate([bps[b], bps[b] - 1]))
            base = np.array([int(rng.choice(bps[k])) for k in range(6)], np.int32)
            A, B = np.meshgrid(va, vb)
            t = np.repeat(base[None, :], A.size, 0)
            t[:, a] = A.ravel()
            t[:, b] = B.ravel()
            rows.append(t)
    tg = np.concatenate(rows).astype(np.int32)
    assert np.array_equal(eng.targets(tg), orc.targets(tg))


def test_nether_targets_and_limits():
    for version in G.VERSIONS:
        eng = mcbiome.Engine(version, "the_nether")
        assert eng.cells()[0] == 0
        orc = O.Oracle(version, "the_nether")
        tg = adversarial_targets(orc, np.random.default_rng(3), 20000)
        tg[:, 2:] = 0
        assert np.array_equal(eng.targets(tg), orc.targets(tg))
    eng = mcbiome.Engine("26.3", "overworld")
    lim = 1 << 24
    tg = np.array([[lim, -lim, lim, -lim, lim, -lim], [-lim, lim, 0, 0, -lim, lim]], np.int32)
    assert np.array_equal(eng.targets(tg), O.Oracle("26.3", "overworld").targets(tg))
    with pytest.raises(mcbiome.EngineError):
        eng.targets(np.array([[lim + 1, 0, 0, 0, 0, 0]], np.int32))
    with pytest.raises(mcbiome.EngineError):
        mcbiome.Engine("26.3", "the_end").targets(np.zeros((1, 6), np.int32))


@pytest.mark.parametrize("version", G.VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether"])
def test_tie_masks_match_oracle(version, dim):
    """mcb_tie_masks and the mcg_ref_tie_mask drop-in (engine/cuda/mcg_host.c:85-96) equal the oracle's
    brute-force tie mask (pinned against the reference engine's on the CPU): adversarial targets on and
    between breakpoints (many ties), targets of real samples, and int32-magnitude targets."""
    eng = mcbiome.Engine(version, dim)
    orc = O.Oracle(version, dim)
    rng = np.random.default_rng(60 + G.VERSIONS.index(version))
    if dim == "overworld":
        tg = adversarial_targets(orc, rng, 4000)
    else:
        tg = rng.integers(-20000, 20000, (4000, 6)).astype(np.int32)
        tg[:, 2:] = 0
    far = rng.integers(-2**30, 2**30, (64, 6)).astype(np.int32)
    tg = np.concatenate([tg, far]).astype(np.int32)
    got = eng.tie_masks(tg)
    for i, t in enumerate(tg):
        _, m = orc.tie_mask(t)
        assert (int(got[i, 0]), int(got[i, 1])) == m, (i, t.tolist())
    host = mcbiome.Host(version, dim)
    for t in tg[:50]:
        assert host.tie_mask(t) == orc.tie_mask(t)[1]
    n_ties = sum(bin(int(a)).count("1") + bin(int(b)).count("1") > 1 for a, b in got)
    print("targets with several tied biomes:", n_ties, "of", len(tg), flush=True)
