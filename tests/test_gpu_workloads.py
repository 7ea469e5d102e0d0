"""The HIP engine at the benchmark workloads' own inputs against the reference CPU engine's outputs
(tests/golden/engine/, engine/cuda/mcg_host.c:36-83 compiled from its sources):
  * the one-seed 256x256 quart grid at quart y 16, through mcb_region and the mcg_ref_points drop-in;
  * all 4096 seeds of the batched workload (64x64 quarts x 4 quart y, origins in +-2^22 blocks), 26.3
    and 26.1, through one plan with the benchmark's launch shape (16-column tiles, 4 columns per lane,
    16,384 tiles): one workgroup per tile, and persistent workgroups that each own 8 tiles, both with
    the XCD-ordered tile ranking active (workgroup counts divisible by 8); per-seed hashes equal."""
import numpy as np
import pytest

from golden_io import engine_fixture, fnv1a64_rows

pytestmark = pytest.mark.gpu
mcbiome = pytest.importorskip("mcbiome")
torch = pytest.importorskip("torch")

BENCH_TILE = 16 | 4 << 16


def test_plain_grid_equals_reference_engine():
    f = engine_fixture("plain_26.3")
    seed, qx0, qz0, qy = int(f["seed"]), int(f["qx0"]), int(f["qz0"]), int(f["qy"])
    eng = mcbiome.Engine("26.3", "overworld")
    got = eng.region(seed, qx0, qz0, 256, 256, [qy])
    assert np.array_equal(got[0], f["grid"])
    xs, zs = np.meshgrid(np.arange(256), np.arange(256))
    pts = np.stack([qx0 + xs.ravel(), np.full(65536, qy), qz0 + zs.ravel()], 1).astype(np.int32)
    h = mcbiome.Host("26.3", "overworld")
    assert np.array_equal(h.ref_points([seed], pts).reshape(256, 256), f["grid"])


@pytest.mark.parametrize("version", ["26.3", "26.1"])
@pytest.mark.parametrize("blocks", [0, 2048])
def test_batched_hashes_equal_reference_engine(version, blocks):
    b = engine_fixture("batched_%s" % version)
    n, qys = len(b["seeds"]), [int(v) for v in b["qys"]]
    per = 64 * 64 * len(qys)
    regs = [mcbiome.make_region(int(b["seeds"][i]), int(b["qx0"][i]), int(b["qz0"][i]), 64, 64, qys, False, i * per)
            for i in range(n)]
    eng = mcbiome.Engine(version, "overworld")
    plan = eng.plan(regs, tile_w=BENCH_TILE, blocks=blocks)
    assert plan.tiles == 16384 and plan.tiles % 8 == 0
    assert blocks % 8 == 0 and blocks < plan.tiles
    assert plan.shape()[3] == (1 if version == "26.3" else 0)        # 26.3: the kernel form without wrap's test
    out = torch.empty(n * per, dtype=torch.uint8, device="cuda")
    out.fill_(0xEE)
    plan.run(out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = fnv1a64_rows(out.cpu().numpy().reshape(n, per))
    bad = np.nonzero(got != b["hashes"])[0]
    assert bad.size == 0, "seeds differing from the reference engine: %s" % bad[:16].tolist()


def test_batched_autotuned_shape_equals_reference_engine():
    """Whatever launch shape the autotuner keeps (the benchmark's), the output stays the reference's."""
    b = engine_fixture("batched_26.3")
    n, qys = 1024, [int(v) for v in b["qys"]]
    per = 64 * 64 * len(qys)
    regs = [mcbiome.make_region(int(b["seeds"][i]), int(b["qx0"][i]), int(b["qz0"][i]), 64, 64, qys, False, i * per)
            for i in range(n)]
    plan = mcbiome.Engine("26.3", "overworld").plan(regs)
    out = torch.empty(n * per, dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    plan.autotune(out.data_ptr(), st)
    out.fill_(0xEE)
    plan.run(out.data_ptr(), st)
    torch.cuda.synchronize()
    assert np.array_equal(fnv1a64_rows(out.cpu().numpy().reshape(n, per)), b["hashes"][:n])
