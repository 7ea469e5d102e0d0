"""Batched plans: many seeds and regions of uneven size in one persistent launch (plus the zoom pass),
eager and hipGraph-replayed, each request's output identical to the same request computed alone and to
the oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
mcbiome = pytest.importorskip("mcbiome")
torch = pytest.importorskip("torch")


def _dev_bytes(n):
    return torch.zeros(int(max(n, 1)), dtype=torch.uint8, device="cuda")


def _mixed_regions(rng, n, zoom_every=0):
    regs, off = [], 0
    for i in range(n):
        seed = int(rng.integers(-2**63, 2**63 - 1, dtype=np.int64))
        side = int(np.round(8 * 8 ** rng.random()))            # 8 .. 64, log-uniform
        x0, z0 = (int(v) for v in rng.integers(-2**20, 2**20, 2))
        zoom = bool(zoom_every and i % zoom_every == 0)
        ys = [64, 70] if zoom else [int(v) for v in rng.integers(-16, 80, int(rng.integers(1, 5)))]
        regs.append(mcbiome.make_region(seed, x0, z0, side, side, ys, zoom, off))
        off += side * side * len(ys)
    return regs, off


@pytest.mark.parametrize("version", ["26.1", "26.3"])
@pytest.mark.parametrize("tile_w,blocks", [(16, 0), (32, 7), (64, 0), (16 | 2 << 16, 0), (32 | 4 << 16, 5), (8 | 4 << 16, 0)])
def test_plan_equals_alone_and_oracle(version, tile_w, blocks):
    rng = np.random.default_rng(41 + tile_w + blocks)
    eng = mcbiome.Engine(version, "overworld")
    orc = O.Oracle(version, "overworld")
    regs, total = _mixed_regions(rng, 12, zoom_every=4)
    plan = eng.plan(regs, tile_w=tile_w, blocks=blocks)
    assert plan.out_bytes == total
    out = _dev_bytes(total)
    plan.run(out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    host = out.cpu().numpy()
    for r in regs:
        ys = [r.y[j] for j in range(r.ny)]
        got = host[r.out_off: r.out_off + r.ny * r.nx * r.nz].reshape(r.ny, r.nz, r.nx)
        alone = eng.region(r.seed, r.x0, r.z0, r.nx, r.nz, ys, zoom=bool(r.zoom))
        assert np.array_equal(got, alone)
        for j, y in enumerate(ys):
            want = orc.block_grid(r.seed, r.x0, r.z0, r.nx, r.nz, 1, y) if r.zoom else \
                orc.grid(r.seed, r.x0, r.z0, r.nx, r.nz, y, 0)
            assert np.array_equal(got[j], want)


def test_plan_graph_replays():
    """hipGraph replay: every replay equals the eager run, with other eager work and syncs in between."""
    rng = np.random.default_rng(5)
    eng = mcbiome.Engine("26.3", "overworld")
    regs, total = _mixed_regions(rng, 10, zoom_every=3)
    plan = eng.plan(regs)
    stream = torch.cuda.current_stream().cuda_stream
    eager = _dev_bytes(total)
    plan.run(eager.data_ptr(), stream)
    torch.cuda.synchronize()
    ref = eager.cpu().numpy()
    out = _dev_bytes(total)
    for rep in range(4):
        out.fill_(0xEE)
        plan.run(out.data_ptr(), stream, graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), ref), "replay %d" % rep
        x = torch.randn(1 << 16, device="cuda")          # unrelated eager work between replays
        (x * 2).sum().item()


def test_points_dev_stream():
    rng = np.random.default_rng(6)
    eng = mcbiome.Engine("26.3", "overworld")
    seeds = rng.integers(-2**63, 2**63 - 1, 40, dtype=np.int64)
    pts = np.stack([rng.integers(-9999, 9999, 70), rng.integers(-16, 80, 70), rng.integers(-9999, 9999, 70)], 1).astype(np.int32)
    d_seeds = torch.from_numpy(seeds).cuda()
    d_pts = torch.from_numpy(pts.ravel()).cuda()
    d_out = torch.zeros(40 * 70, dtype=torch.uint8, device="cuda")
    d_tg = torch.zeros(40 * 70 * 6, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        eng.points_dev(d_seeds.data_ptr(), 40, d_pts.data_ptr(), 70, d_out.data_ptr(), d_tg.data_ptr(), s.cuda_stream)
    s.synchronize()
    want, wtg = O.Oracle("26.3", "overworld").points(seeds, pts, want_tg=True)
    assert np.array_equal(d_out.cpu().numpy().reshape(40, 70), want)
    assert np.array_equal(d_tg.cpu().numpy().reshape(40, 70, 6), wtg)


def test_empty_and_bad_inputs():
    eng = mcbiome.Engine("26.3", "overworld")
    assert eng.points(np.zeros(0, np.int64), np.zeros((3, 3), np.int32)).size == 0
    with pytest.raises(mcbiome.EngineError):
        eng.plan([mcbiome.make_region(1, 0, 0, 0, 4, [16])])
    with pytest.raises(mcbiome.EngineError):
        eng.plan([mcbiome.make_region(1, 0, 0, 4, 4, list(range(9)))])


@pytest.mark.parametrize("version,dim", [("26.3", "overworld"), ("26.1", "overworld"), ("26.1", "the_nether")])
def test_wrap_free_kernel_choice_and_parity(version, dim):
    """A grid runs without PerlinNoise.wrap's range test only when every octave input of the launch
    provably stays inside +-2^24 (quart coordinates up to about 8e6 in the Overworld, 26.3 and, up to two
    layers per column, 26.1/26.2); plans reaching beyond keep the test.  Results equal the oracle on both
    sides of the threshold, in one plan and mixed (one far request puts the whole launch on the tested
    form)."""
    eng = mcbiome.Engine(version, dim)
    orc = O.Oracle(version, dim)
    near = [(11, -2**20, 2**20 - 40), (-5, 8_000_000 - 64, -8_000_000)]            # inside: no test needed
    far = [(77, 2**26 - 30, -(2**24) - 7), (1, -(2**27) + 5, 2**25 + 3)]           # wrap decides rounding
    # the Nether's highest octave frequency is 2^-6: its inputs reach 2^24 only near quart 2^29
    edge = far if dim == "overworld" else [(3, 2**29 - 64, -(2**29) + 8)]
    checks = [(near, True), (edge, False), (near + edge, False)] + ([] if dim == "overworld" else [(far, True)])
    for cases, expect_free in checks:
        regs, off = [], 0
        for seed, x0, z0 in cases:
            regs.append(mcbiome.make_region(seed, x0, z0, 40, 24, [16, -4], False, off))
            off += 40 * 24 * 2
        plan = eng.plan(regs, tile_w=16 | 2 << 16)
        assert plan.shape()[3] == (1 if expect_free else 0), cases
        out = _dev_bytes(off)
        plan.run(out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        host = out.cpu().numpy()
        for r in regs:
            got = host[r.out_off: r.out_off + 2 * 40 * 24].reshape(2, 24, 40)
            for j, y in enumerate([16, -4]):
                assert np.array_equal(got[j], orc.grid(r.seed, r.x0, r.z0, 40, 24, y, 0)), (r.seed, r.x0, j)
