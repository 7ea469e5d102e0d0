"""Mixed-world plans, the autotuner and the continuous-batching server: every request's result equals
the same request computed alone (and the oracle), whatever batch it lands in."""
import threading

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
mcbiome = pytest.importorskip("mcbiome")
torch = pytest.importorskip("torch")

WORLDS = [("26.3", "overworld"), ("26.1", "overworld"), ("26.2", "the_nether"), ("26.3", "the_end"),
          ("26.2", "overworld")]


def _requests(rng, ctx, n, side_lo=8, side_hi=48, seeds=None):
    widx = [ctx.world(v, d) for v, d in WORLDS]
    regs, meta, off = [], [], 0
    for i in range(n):
        k = int(rng.integers(0, len(WORLDS)))
        v, d = WORLDS[k]
        seed = int(seeds[i % len(seeds)]) if seeds is not None else int(rng.integers(-2**63, 2**63 - 1, dtype=np.int64))
        side = int(rng.integers(side_lo, side_hi))
        if d == "the_end":
            side = min(side, 20)
        zoom = i % 5 == 0
        ys = [64, -30][: 1 + (i % 2)] if zoom else [int(v) for v in rng.integers(-16, 80, 1 + (i % 3))]
        r = mcbiome.make_region(seed, int(rng.integers(-2**19, 2**19)), int(rng.integers(-2**19, 2**19)), side, side, ys,
                                zoom, off)
        r.world = widx[k]
        regs.append(r)
        meta.append((v, d))
        off += side * side * len(ys)
    return regs, meta, off


def _alone(engines, v, d, r):
    key = (v, d)
    if key not in engines:
        engines[key] = mcbiome.Engine(v, d)
    return engines[key].region(r.seed, r.x0, r.z0, r.nx, r.nz, [r.y[j] for j in range(r.ny)], zoom=bool(r.zoom))


def test_context_plan_mixed_worlds_and_autotune():
    rng = np.random.default_rng(12)
    ctx = mcbiome.Context()
    regs, meta, total = _requests(rng, ctx, 25)
    plan = ctx.plan(regs)
    out = torch.zeros(total, dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    plan.run(out.data_ptr(), st)
    torch.cuda.synchronize()
    first = out.cpu().numpy().copy()
    engines = {}
    for r, (v, d) in zip(regs, meta):
        got = first[r.out_off: r.out_off + r.ny * r.nx * r.nz].reshape(r.ny, r.nz, r.nx)
        assert np.array_equal(got, _alone(engines, v, d, r)), (v, d, r.seed)
    tw, ms = plan.autotune(out.data_ptr(), st)
    assert (tw & 0xFFFF) in (8, 16, 32, 64) and (tw >> 16) in (1, 2, 4, 16) and ms > 0
    out.zero_()
    plan.run(out.data_ptr(), st)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), first), "autotuned launch shape changed results"
    # spot-check against the oracle
    for r, (v, d) in list(zip(regs, meta))[:6]:
        got = first[r.out_off: r.out_off + r.ny * r.nx * r.nz].reshape(r.ny, r.nz, r.nx)
        o = O.Oracle(v, d)
        for j in range(r.ny):
            want = o.block_grid(r.seed, r.x0, r.z0, r.nx, r.nz, 1, r.y[j]) if r.zoom else \
                o.grid(r.seed, r.x0, r.z0, r.nx, r.nz, r.y[j], 0)
            assert np.array_equal(got[j], want)


@pytest.mark.parametrize("batch_bytes,lanes", [(1 << 16, 2), (8 << 20, 3)])
def test_server_results_equal_alone(batch_bytes, lanes):
    """Requests arriving while others run are admitted into later batches; repeated seeds hit the
    device slot cache; each result equals the request computed alone."""
    rng = np.random.default_rng(batch_bytes + lanes)
    ctx = mcbiome.Context()
    seeds = rng.integers(-2**63, 2**63 - 1, 7, dtype=np.int64)      # few seeds: cache hits across batches
    regs, meta, _ = _requests(rng, ctx, 60, seeds=seeds)
    srv = mcbiome.Server(ctx, batch_bytes=batch_bytes, lanes=lanes, slot_capacity=64)
    bufs = [np.zeros(r.ny * r.nx * r.nz, np.uint8) for r in regs]
    tickets = [None] * len(regs)

    def submit(lo, hi):
        for i in range(lo, hi):
            tickets[i] = srv.submit(regs[i], bufs[i])

    th = [threading.Thread(target=submit, args=(k * 20, (k + 1) * 20)) for k in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    lat = [srv.wait(t) for t in tickets]
    assert all(x >= 0 for x in lat)
    stats = srv.stats()
    assert stats["completed"] == len(regs)
    assert stats["batches"] >= 2
    assert stats["slot_hit_rate"] > 0.3
    engines = {}
    for r, (v, d), b in zip(regs, meta, bufs):
        assert np.array_equal(b.reshape(r.ny, r.nz, r.nx), _alone(engines, v, d, r)), (v, d, r.seed)
    srv.close()
    ctx.close()


def test_server_rejects_oversized_request():
    ctx = mcbiome.Context()
    w = ctx.world("26.3", "overworld")
    srv = mcbiome.Server(ctx, batch_bytes=1 << 12, lanes=1, slot_capacity=8)
    r = mcbiome.make_region(1, 0, 0, 128, 128, [16])
    r.world = w
    with pytest.raises(mcbiome.EngineError):
        srv.submit(r, np.zeros(128 * 128, np.uint8))
    srv.close()


def test_server_slot_pressure_defers_and_stays_exact():
    """More distinct seeds than seed slots: admission defers requests until slots are unpinned instead of
    failing a batch, so no seed ever maps to a slot whose state was not written for it.  Each request is
    submitted twice (the second round hits or re-initialises slots evicted meanwhile); every result
    equals the request computed alone."""
    rng = np.random.default_rng(77)
    ctx = mcbiome.Context()
    w = ctx.world("26.3", "overworld")
    seeds = rng.integers(-2**63, 2**63 - 1, 24, dtype=np.int64)
    regs = []
    for i, s in enumerate(seeds):
        r = mcbiome.make_region(int(s), int(rng.integers(-2**18, 2**18)), int(rng.integers(-2**18, 2**18)), 16, 16,
                                [16, -8][: 1 + i % 2])
        r.world = w
        regs.append(r)
    # a batch arena of 64 KiB holds every request at once; 5 slots hold 5 of the 24 seeds
    srv = mcbiome.Server(ctx, batch_bytes=1 << 16, lanes=3, slot_capacity=5)
    engines = {}
    want = [_alone(engines, "26.3", "overworld", r) for r in regs]
    for rnd in range(2):
        bufs = [np.zeros(r.ny * r.nx * r.nz, np.uint8) for r in regs]
        tickets = [srv.submit(r, b) for r, b in zip(regs, bufs)]
        for t in tickets:
            srv.wait(t)
        for i, (r, b) in enumerate(zip(regs, bufs)):
            assert np.array_equal(b.reshape(r.ny, r.nz, r.nx), want[i]), ("round", rnd, "request", i)
    assert srv.stats()["batches"] >= 2 * (24 // 5)
    srv.close()
    ctx.close()
