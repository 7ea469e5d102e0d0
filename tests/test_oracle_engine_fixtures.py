"""The oracle against the reference engine's own outputs at the benchmark workloads' inputs
(tests/golden/engine/, written from engine/cuda/mcg_host.c:36-83 compiled from its sources): the
one-seed 256x256 grid in full, and per-seed hashes of a sample of the 4096-seed batch."""
import numpy as np
import pytest

import oracle as O
from golden_io import engine_fixture, fnv1a64_rows, signed64, splitmix64, uniform


def _points(qx0, qz0, nx, nz, qys):
    xs, zs = np.meshgrid(np.arange(nx), np.arange(nz))
    return np.concatenate([np.stack([(qx0 + xs).ravel(), np.full(nx * nz, qy), (qz0 + zs).ravel()], 1)
                           for qy in qys]).astype(np.int32)


def test_fixture_inputs_follow_the_stream():
    """The stored seeds and origins are the splitmix64 draws the workloads name."""
    f = engine_fixture("plain_26.3")
    d = splitmix64(0, 3)
    assert int(f["seed"]) == signed64(d[0])
    assert int(f["qx0"]) == uniform(d[1], 1 << 20) >> 2 and int(f["qz0"]) == uniform(d[2], 1 << 20) >> 2
    b = engine_fixture("batched_26.3")
    d = splitmix64(1, 3 * 4096)
    assert b["seeds"].tolist() == [signed64(d[3 * i]) for i in range(4096)]
    assert b["qx0"].tolist() == [uniform(d[3 * i + 1], 1 << 22) >> 2 for i in range(4096)]
    assert b["qz0"].tolist() == [uniform(d[3 * i + 2], 1 << 22) >> 2 for i in range(4096)]
    assert np.array_equal(b["hashes"][:8], fnv1a64_rows(b["first_grids"].reshape(8, -1)))


def test_oracle_plain_grid_equals_reference_engine():
    f = engine_fixture("plain_26.3")
    got = O.Oracle("26.3", "overworld").grid(int(f["seed"]), int(f["qx0"]), int(f["qz0"]), 256, 256, int(f["qy"]), 0)
    assert np.array_equal(got, f["grid"])


@pytest.mark.parametrize("version", ["26.3", "26.1"])
def test_oracle_batched_sample_equals_reference_engine(version):
    b = engine_fixture("batched_%s" % version)
    orc = O.Oracle(version, "overworld")
    idx = list(range(0, 4096, 128)) + [4095]
    rows = np.stack([orc.points([int(b["seeds"][i])], _points(int(b["qx0"][i]), int(b["qz0"][i]), 64, 64, b["qys"])).ravel()
                     for i in idx])
    assert np.array_equal(fnv1a64_rows(rows), b["hashes"][idx])
