#!/usr/bin/env python3
"""Write the reference-engine fixtures of the benchmark workloads into tests/golden/engine/.

Runs only in a container that holds the reference sources: the reference CPU engine
(engine/cuda/mcg_host.c:36-83, `mcg_ref_points`) compiled from them by `make -C oracle ref` into
oracle/_ref/libmcg_ref.so.  Its outputs are committed as data, so the GPU tests need neither the
reference nor its build.

Workloads (seeds and origins from the splitmix64 stream, state0 = k, draws as unsigned 64-bit ints;
uniform in [-R, R] = draw mod (2R+1) - R; quart = block >> 2):
  plain_26.3.npz   one seed = draw 0 of splitmix64(0); origin blocks = draws 1, 2 in +-2^20;
                   256x256 quarts at quart y 16, Overworld 26.3 -> the whole u8 grid [iz][ix].
  batched_<V>.npz  4096 seeds, draws 3i, 3i+1, 3i+2 of splitmix64(1) = seed, block x0, block z0 in
                   +-2^22; each a 64x64 quart grid at quart y {-16, 0, 16, 48}; Overworld 26.3 and 26.1
                   -> per seed the FNV-1a-64 hash of its 16,384 output bytes in region order
                   (layer, iz, ix), plus the full grids of the first 8 seeds.
Usage: python tests/golden/make_engine_fixtures.py   (about two minutes on 8 host threads)
"""
import os
import sys
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from golden_io import fnv1a64_rows, splitmix64, uniform, signed64  # noqa: E402

OUT = os.path.join(HERE, "engine")
QYS = (-16, 0, 16, 48)


def region_points(qx0, qz0, nx, nz, qys):
    xs, zs = np.meshgrid(np.arange(nx), np.arange(nz))
    return np.concatenate([np.stack([(qx0 + xs).ravel(), np.full(nx * nz, qy), (qz0 + zs).ravel()], 1)
                           for qy in qys]).astype(np.int32)


def plain():
    d = splitmix64(0, 3)
    seed, bx0, bz0 = signed64(d[0]), uniform(d[1], 1 << 20), uniform(d[2], 1 << 20)
    ref = oracle.RefEngine("26.3", "overworld")
    grid = ref.points([seed], region_points(bx0 >> 2, bz0 >> 2, 256, 256, [16])).reshape(256, 256)
    np.savez_compressed(os.path.join(OUT, "plain_26.3.npz"), seed=np.int64(seed), qx0=np.int32(bx0 >> 2),
                        qz0=np.int32(bz0 >> 2), qy=np.int32(16), grid=grid)
    print("plain: seed %d origin quart (%d, %d) hash %016x" % (seed, bx0 >> 2, bz0 >> 2, fnv1a64_rows(grid.reshape(1, -1))[0]))


def batched(version, nthreads=8):
    n = 4096
    d = splitmix64(1, 3 * n)
    work = [(signed64(d[3 * i]), uniform(d[3 * i + 1], 1 << 22) >> 2, uniform(d[3 * i + 2], 1 << 22) >> 2) for i in range(n)]
    ref = oracle.RefEngine(version, "overworld")
    out = np.zeros((n, 64 * 64 * len(QYS)), np.uint8)
    state = {"next": 0}
    lock = threading.Lock()

    def worker():
        while True:
            with lock:
                i = state["next"]
                state["next"] += 1
            if i >= n:
                return
            s, qx0, qz0 = work[i]
            out[i] = ref.points([s], region_points(qx0, qz0, 64, 64, QYS)).ravel()

    th = [threading.Thread(target=worker) for _ in range(nthreads)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    hashes = fnv1a64_rows(out)
    np.savez_compressed(os.path.join(OUT, "batched_%s.npz" % version),
                        seeds=np.array([w[0] for w in work], np.int64),
                        qx0=np.array([w[1] for w in work], np.int32), qz0=np.array([w[2] for w in work], np.int32),
                        qys=np.array(QYS, np.int32), hashes=hashes, first_grids=out[:8].reshape(8, len(QYS), 64, 64))
    print("batched %s: %d seeds, hash of hashes %016x" % (version, n, fnv1a64_rows(hashes.view(np.uint8).reshape(1, -1))[0]))


if __name__ == "__main__":
    if not oracle.RefEngine.available():
        sys.exit("oracle/_ref/libmcg_ref.so missing: build it with `make -C oracle ref` (needs the reference sources)")
    os.makedirs(OUT, exist_ok=True)
    plain()
    for v in ("26.3", "26.1"):
        batched(v)
