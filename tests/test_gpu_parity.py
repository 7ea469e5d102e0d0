"""Parity of the HIP engine (libmcbiome_hip.so, through its C ABI) with the oracle: bit-exact biome ids
and climate targets on the same seeded inputs, every version and dimension."""
import numpy as np
import pytest

import golden_io as G
import oracle as O

pytestmark = pytest.mark.gpu

mcbiome = pytest.importorskip("mcbiome")
NAMES = O.biome_names()
EDGE_SEEDS = [0, 1, -1, 2**63 - 1, -2**63, 12345, 8675309, -7048155917072976836]


@pytest.fixture(scope="module", autouse=True)
def _device():
    if mcbiome.device_count() < 1:
        pytest.fail("no HIP device visible to libmcbiome_hip.so")


def _rand_points(rng, n, span, qy_lo=-16, qy_hi=80):
    return np.stack([rng.integers(-span, span, n), rng.integers(qy_lo, qy_hi, n), rng.integers(-span, span, n)],
                    1).astype(np.int32)


@pytest.mark.parametrize("version", G.VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether", "the_end"])
def test_points_match_oracle(version, dim):
    rng = np.random.default_rng(100 + G.VERSIONS.index(version) * 3 + ["overworld", "the_nether", "the_end"].index(dim))
    seeds = np.array(EDGE_SEEDS + [int(v) for v in rng.integers(-2**63, 2**63 - 1, 24, dtype=np.int64)], np.int64)
    npts = 96 if dim != "the_end" else 24
    pts = np.concatenate([_rand_points(rng, npts // 3, 2**8), _rand_points(rng, npts // 3, 2**18),
                          _rand_points(rng, npts - 2 * (npts // 3), 2**20)])
    print("points", version, dim, seeds.size, "x", pts.shape[0], flush=True)
    want, want_tg = O.Oracle(version, dim).points(seeds, pts, want_tg=True)
    got, got_tg = mcbiome.Engine(version, dim).points(seeds, pts, want_tg=True)
    assert np.array_equal(got_tg, want_tg), "climate targets differ"
    assert np.array_equal(got, want), "biomes differ at %d samples" % int((got != want).sum())


@pytest.mark.parametrize("version", G.VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether", "the_end"])
def test_points_few_per_seed(version, dim):
    """Few points per seed take the kernel that reads each seed's tables from its slot (one lane per seed
    and point: below 128 points per seed in the Overworld, 32 in the Nether, 256 in the End); more stage
    the seed in LDS.  Both equal the oracle on the same inputs, on each side of each threshold."""
    rng = np.random.default_rng(300 + G.VERSIONS.index(version) * 3 + ["overworld", "the_nether", "the_end"].index(dim))
    seeds = np.array(EDGE_SEEDS + [int(v) for v in rng.integers(-2**63, 2**63 - 1, 56, dtype=np.int64)], np.int64)
    span = 2**12 if dim == "the_end" else 2**18
    pts = np.concatenate([_rand_points(rng, 20, span), _rand_points(rng, 20, 2**8), _rand_points(rng, 217, span)])
    print("few points", version, dim, seeds.size, flush=True)
    want, want_tg = O.Oracle(version, dim).points(seeds, pts, want_tg=True)
    eng = mcbiome.Engine(version, dim)
    for k in (1, 7, 31, 32, 40, 127, 128, 255, 257):
        got, got_tg = eng.points(seeds, pts[:k], want_tg=True)
        assert np.array_equal(got_tg, want_tg[:, :k]), "targets differ at %d points per seed" % k
        assert np.array_equal(got, want[:, :k]), "biomes differ at %d points per seed" % k


@pytest.mark.parametrize("version", G.VERSIONS)
def test_refpoints_dropin(version):
    """mcg_host_load + mcg_ref_points (the reference's CPU-engine entry points) served by the HIP engine."""
    rng = np.random.default_rng(7)
    seeds = np.array(EDGE_SEEDS, np.int64)
    pts = _rand_points(rng, 64, 2**16)
    h = mcbiome.Host(version, "overworld")
    assert h.mode == (1 if version == "26.3" else 0)
    got, got_tg = h.ref_points(seeds, pts, want_tg=True)
    want, want_tg = O.Oracle(version, "overworld").points(seeds, pts, want_tg=True)
    assert np.array_equal(got, want) and np.array_equal(got_tg, want_tg)
    h.close()


@pytest.mark.parametrize("version", G.VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether"])
def test_golden_climate_vectors(version, dim):
    """The reference's Java-oracle climate vectors through the HIP engine: exact targets; biomes exact
    except at brute-force-verified fitness ties (the Java resolver keeps search history)."""
    rows = G.climate_rows(version, dim)
    eng = mcbiome.Engine(version, dim)
    orc = O.Oracle(version, dim)
    bad_tg = bad_b = 0
    for seed, rs in G.by_seed(rows).items():
        pts = np.array([[r[1], r[2], r[3]] for r in rs], np.int32)
        b, tg = eng.points([seed], pts, want_tg=True)
        for i, r in enumerate(rs):
            bad_tg += list(tg[0, i]) != r[4]
            if NAMES[b[0, i]] != r[5]:
                cnt, m = orc.tie_mask(r[4])
                bi = int(b[0, i])
                if not (cnt >= 2 and (m[bi >> 6] >> (bi & 63)) & 1):
                    bad_b += 1
    assert bad_tg == 0 and bad_b == 0


@pytest.mark.parametrize("version", G.VERSIONS)
def test_golden_end_vectors(version):
    rows = G.climate_rows(version, "the_end")
    eng = mcbiome.Engine(version, "the_end")
    bad = 0
    for seed, rs in G.by_seed(rows).items():
        pts = np.array([[r[1], r[2], r[3]] for r in rs], np.int32)
        b = eng.points([seed], pts)
        bad += sum(NAMES[b[0, i]] != r[6] for i, r in enumerate(rs))
    assert bad == 0


@pytest.mark.parametrize("version", G.VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether", "the_end"])
def test_region_grid_matches_oracle(version, dim):
    """Quart grids with several y layers (column climate shared across layers) = per-sample oracle."""
    eng = mcbiome.Engine(version, dim)
    orc = O.Oracle(version, dim)
    cases = [(12345, -37, 91, 70, 45, [-16, 0, 16, 48]), (-1, 250000, -180000, 33, 64, [16]),
             (2**63 - 1, -(2**20), 2**20 - 40, 40, 17, [5, 79])]
    if dim == "the_end":
        cases = [(s, x0, z0, min(nx, 24), min(nz, 24), ys) for s, x0, z0, nx, nz, ys in cases]
    for seed, x0, z0, nx, nz, ys in cases:
        print("region", version, dim, seed, x0, z0, nx, nz, ys, flush=True)
        got = eng.region(seed, x0, z0, nx, nz, ys)
        for j, qy in enumerate(ys):
            want = orc.grid(seed, x0, z0, nx, nz, qy, semantics=0)
            assert np.array_equal(got[j], want), (seed, qy, int((got[j] != want).sum()))


@pytest.mark.parametrize("version", G.VERSIONS)
@pytest.mark.parametrize("dim", ["overworld", "the_nether", "the_end"])
def test_zoom_region_matches_oracle(version, dim):
    """Block-resolution grids through BiomeManager zoom = oracle block grid."""
    eng = mcbiome.Engine(version, dim)
    orc = O.Oracle(version, dim)
    n = 40 if dim != "the_end" else 20
    for seed, x0, z0, ys in [(0, -21, 13, [64, -32, 128, 256]), (8675309, 1 << 20, -(1 << 19) + 3, [70])]:
        print("zoom", version, dim, seed, x0, z0, ys, flush=True)
        got = eng.region(seed, x0, z0, n, n, ys, zoom=True)
        for j, y in enumerate(ys):
            want = orc.block_grid(seed, x0, z0, n, n, 1, y)
            assert np.array_equal(got[j], want), (seed, y, int((got[j] != want).sum()))


def test_blockbiome_vectors():
    """misc_biome.jsonl blockbiome records (Java oracle) through the zoomed region path."""
    for version in G.VERSIONS:
        engines = {}
        for r in G.misc_records(version, "blockbiome"):
            key = (r["dim"], r.get("preset", "normal"))
            if key not in engines:
                engines[key] = mcbiome.Engine(version, r["dim"], r.get("preset", "normal"))
            b = engines[key].region(r["seed"], r["x"], r["z"], 1, 1, [r["y"]], zoom=True)[0, 0, 0]
            assert NAMES[b] == r["biome"].replace("minecraft:", ""), r


@pytest.mark.parametrize("preset", ["amplified", "large_biomes"])
def test_presets(preset):
    rng = np.random.default_rng(9)
    seeds = np.array(EDGE_SEEDS[:4], np.int64)
    pts = _rand_points(rng, 64, 2**18)
    for version in G.VERSIONS:
        want, wt = O.Oracle(version, "overworld", preset).points(seeds, pts, want_tg=True)
        got, gt = mcbiome.Engine(version, "overworld", preset).points(seeds, pts, want_tg=True)
        assert np.array_equal(got, want) and np.array_equal(gt, wt)


def test_wrap_branch_far_coordinates():
    """|x*f| >= 2^24 takes PerlinNoise.wrap's outer branch: quart coordinates near 2^27 reach it for the
    highest-frequency octaves.  GPU and oracle must agree (this branch is pinned by closed form only)."""
    seeds = np.array([0, 77], np.int64)
    pts = np.array([[2**27 - 5, 16, -(2**27) + 3], [-(2**27) + 11, 0, 2**27 - 1], [2**26 + 7, -8, 3]], np.int32)
    for version in G.VERSIONS:
        want, wt = O.Oracle(version, "overworld").points(seeds, pts, want_tg=True)
        got, gt = mcbiome.Engine(version, "overworld").points(seeds, pts, want_tg=True)
        assert np.array_equal(gt, wt) and np.array_equal(got, want)


@pytest.mark.parametrize("version", G.VERSIONS)
def test_end_tiles_match_oracle(version):
    """End tiles (island cells listed once per tile) over the centre boundary, negative chunks and far
    coordinates, for several tile shapes and columns per lane."""
    torch = pytest.importorskip("torch")
    eng = mcbiome.Engine(version, "the_end")
    orc = O.Oracle(version, "the_end")
    cases = [(1, -300, -290, 70, 50), (-7048155917072976836, 250, -20, 40, 40), (2**63 - 1, -(2**20), 2**20 - 99, 37, 61)]
    for tile in (16, 64 | 4 << 16, 8 | 16 << 16):
        regs, off = [], 0
        for seed, x0, z0, nx, nz in cases:
            regs.append(mcbiome.make_region(seed, x0, z0, nx, nz, [16], False, off))
            off += nx * nz
        plan = eng.plan(regs, tile_w=tile)
        out = torch.zeros(off, dtype=torch.uint8, device="cuda")
        plan.run(out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        host = out.cpu().numpy()
        for r, (seed, x0, z0, nx, nz) in zip(regs, cases):
            want = orc.grid(seed, x0, z0, nx, nz, 16, 0)
            got = host[r.out_off: r.out_off + nx * nz].reshape(nz, nx)
            assert np.array_equal(got, want), (tile, seed, int((got != want).sum()))


@pytest.mark.parametrize("version", G.VERSIONS)
def test_end_field_matches_oracle(version):
    """mcb_end_field (End island height value and erosion over block grids, many seeds at once) = oracle,
    bit for bit: near the centre island, negative coordinates, fine and coarse steps (a step of 1024
    blocks exceeds the tile's listed window and takes the direct path), far coordinates."""
    eng = mcbiome.Engine(version, "the_end")
    orc = O.Oracle(version, "the_end")
    seeds = np.array(EDGE_SEEDS[:5], np.int64)
    for bx0, bz0, nx, nz, step in [(-700, -650, 40, 37, 8), (5000, -9000, 33, 20, 1), (-2**20, 2**20 - 4096, 18, 17, 37),
                                   (-20000, -20000, 20, 20, 1024), (123456, 654321, 16, 16, 8)]:
        for kind in (0, 1):
            print("end_field", version, bx0, bz0, nx, nz, step, kind, flush=True)
            got = eng.end_field(seeds, bx0, bz0, nx, nz, step, kind)
            for i, s in enumerate(seeds):
                want = orc.end_field(int(s), bx0, bz0, nx, nz, step, kind)
                assert np.array_equal(got[i].view(np.uint64), want.view(np.uint64)), (bx0, step, kind, int(s))
    with pytest.raises(mcbiome.EngineError):
        mcbiome.Engine(version, "overworld").end_field(seeds, 0, 0, 4, 4)
