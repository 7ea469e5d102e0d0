"""mcb_points_dev_noalloc (the points path on a static device scratch array, no allocation) through the C
ABI: bit-identical to mcb_points_dev on the same device buffers, on two calls in a row (the second
smaller, so stale state from the first must not leak), in every dimension; sizes beyond its stated
maximum are refused."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
mcbiome = pytest.importorskip("mcbiome")
torch = pytest.importorskip("torch")


def _call(fn, eng, d_seeds, ns, d_pts, np_, d_out, d_tg, stream):
    return fn(eng.h, C.c_void_p(d_seeds.data_ptr()), ns, C.c_void_p(d_pts.data_ptr()), np_,
              C.c_void_p(d_out.data_ptr()), C.c_void_p(d_tg.data_ptr() if d_tg is not None else None), C.c_void_p(stream))


@pytest.mark.parametrize("version,dim", [("26.3", "overworld"), ("26.1", "overworld"), ("26.2", "the_nether"),
                                         ("26.3", "the_end")])
def test_noalloc_equals_points_dev(version, dim):
    L = mcbiome.lib()
    eng = mcbiome.Engine(version, dim)
    rng = np.random.default_rng(zlib_seed(version, dim))
    stream = torch.cuda.current_stream().cuda_stream
    for ns, np_ in ((1024, 300), (37, 9)):                      # second call smaller
        seeds = rng.integers(-2**63, 2**63 - 1, ns, dtype=np.int64)
        seeds[:3] = [0, 1, -1]
        pts = np.stack([rng.integers(-2**20, 2**20, np_), rng.integers(-16, 80, np_),
                        rng.integers(-2**20, 2**20, np_)], 1).astype(np.int32)
        d_seeds = torch.from_numpy(seeds).cuda()
        d_pts = torch.from_numpy(pts.ravel()).cuda()
        outs = []
        for fn in (L.mcb_points_dev, L.mcb_points_dev_noalloc):
            d_out = torch.full((ns * np_,), 0xEE, dtype=torch.uint8, device="cuda")
            d_tg = torch.full((ns * np_ * 6,), -7, dtype=torch.int32, device="cuda")
            assert _call(fn, eng, d_seeds, ns, d_pts, np_, d_out, d_tg, stream) == 0, L.mcb_last_error()
            torch.cuda.synchronize()
            outs.append((d_out.cpu().numpy(), d_tg.cpu().numpy()))
        assert np.array_equal(outs[0][0], outs[1][0]), (ns, np_)
        assert np.array_equal(outs[0][1], outs[1][1]), (ns, np_)
        k = min(ns, 8)
        want = O.Oracle(version, dim).points(seeds[:k], pts)
        assert np.array_equal(outs[1][0].reshape(ns, np_)[:k], want)


def test_noalloc_rejects_beyond_max():
    L = mcbiome.lib()
    eng = mcbiome.Engine("26.3", "overworld")
    n = 1025
    d_seeds = torch.zeros(n, dtype=torch.int64, device="cuda")
    d_pts = torch.zeros(3, dtype=torch.int32, device="cuda")
    d_out = torch.zeros(n, dtype=torch.uint8, device="cuda")
    rc = _call(L.mcb_points_dev_noalloc, eng, d_seeds, n, d_pts, 1, d_out, None, 0)
    assert rc == 6 and b"at most 1024" in L.mcb_last_error()


def zlib_seed(version, dim):
    import zlib
    return zlib.crc32(("noalloc/%s/%s" % (version, dim)).encode())
