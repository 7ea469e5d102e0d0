"""26.4-snapshot-2 on the device (engine/mc_common.h:38): the 26.3 climate with two changed parameter rows
and exclusive upper parameter bounds (engine/mc_data.h:178-180), searched by the exact tree walk (no cell
table).  Points, targets, tie masks, quart grids and zoomed block grids of a batched plan, and the
mcg_ref_points drop-in equal the oracle, whose 26.4 is pinned against the reference engine on the CPU."""
import numpy as np
import pytest

import oracle as O
from test_oracle_targets import adversarial_targets

pytestmark = pytest.mark.gpu
mcbiome = pytest.importorskip("mcbiome")
torch = pytest.importorskip("torch")
V = "26.4-snapshot-2"
EDGE_SEEDS = [0, 1, -1, 2**63 - 1, -2**63, 12345]


@pytest.mark.parametrize("dim", ["overworld", "the_nether", "the_end"])
def test_points_match_oracle(dim):
    rng = np.random.default_rng(264 + ["overworld", "the_nether", "the_end"].index(dim))
    seeds = np.array(EDGE_SEEDS + [int(v) for v in rng.integers(-2**63, 2**63 - 1, 26, dtype=np.int64)], np.int64)
    n = 300 if dim != "the_end" else 40
    pts = np.stack([rng.integers(-2**20, 2**20, n), rng.integers(-16, 80, n), rng.integers(-2**20, 2**20, n)], 1).astype(np.int32)
    want, want_tg = O.Oracle(V, dim).points(seeds, pts, want_tg=True)
    eng = mcbiome.Engine(V, dim)
    for k in (n, 20):                                 # staged and slot-reading kernels
        got, got_tg = eng.points(seeds, pts[:k], want_tg=True)
        assert np.array_equal(got_tg, want_tg[:, :k]) and np.array_equal(got, want[:, :k]), k
    h = mcbiome.Host(V, dim)
    assert np.array_equal(h.ref_points(seeds[:4], pts), want[:4])


@pytest.mark.parametrize("dim", ["overworld", "the_nether"])
def test_targets_and_tie_masks_match_oracle(dim):
    eng = mcbiome.Engine(V, dim)
    assert eng.cells()[0] == 0                        # the walk serves every search
    orc = O.Oracle(V, dim)
    rng = np.random.default_rng(2641)
    tg = adversarial_targets(orc, rng, 20000) if dim == "overworld" else rng.integers(-20000, 20000, (20000, 6)).astype(np.int32)
    if dim != "overworld":
        tg[:, 2:] = 0
    assert np.array_equal(eng.targets(tg), orc.targets(tg))
    m = eng.tie_masks(tg[:2000])
    for i in range(0, 2000, 7):
        assert (int(m[i, 0]), int(m[i, 1])) == orc.tie_mask(tg[i])[1]


def test_plan_grids_match_oracle():
    rng = np.random.default_rng(2642)
    eng = mcbiome.Engine(V, "overworld")
    orc = O.Oracle(V, "overworld")
    regs, off = [], 0
    for i in range(10):
        side = int(rng.integers(16, 48))
        zoom = i % 3 == 0
        ys = [64] if zoom else [int(v) for v in rng.integers(-16, 80, 2)]
        regs.append(mcbiome.make_region(int(rng.integers(-2**63, 2**63 - 1, dtype=np.int64)), int(rng.integers(-2**20, 2**20)),
                                        int(rng.integers(-2**20, 2**20)), side, side, ys, zoom, off))
        off += side * side * len(ys)
    plan = eng.plan(regs, tile_w=16 | 4 << 16)
    out = torch.zeros(off, dtype=torch.uint8, device="cuda")
    plan.run(out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    host = out.cpu().numpy()
    for r in regs:
        got = host[r.out_off: r.out_off + r.ny * r.nx * r.nz].reshape(r.ny, r.nz, r.nx)
        for j in range(r.ny):
            want = orc.block_grid(r.seed, r.x0, r.z0, r.nx, r.nz, 1, r.y[j]) if r.zoom else \
                orc.grid(r.seed, r.x0, r.z0, r.nx, r.nz, r.y[j], 0)
            assert np.array_equal(got[j], want), (r.seed, j)


def test_plugin_forms_refuse_the_snapshot():
    eng = mcbiome.Engine(V, "overworld")
    with pytest.raises(mcbiome.EngineError):
        eng.plugin_world(1)
