"""The large 3D slope (16,806,912 particles on a 256 x 128 x 128 hex8 mesh,
4,276,737 nodes) at full size, through properties that do not need the
oracle to run at that size: after 4 steps every particle's cell is the
structured-mesh cell that contains its position (the binning is exact) and
its local coordinate is 2 (x - centroid) / h; and the fast mode (cell-sorted
scatter with FP64 atomics) agrees with the ordered mode (particle-order sums,
bit-identical to the oracle at the sizes the oracle runs) to 1e-12 relative
per step."""
import numpy as np
import pytest

from mpm_amd import workloads

pytestmark = pytest.mark.gpu

STEPS = 4


def _run(ordered):
    from mpm_amd.capi import Solver
    w = workloads.wedge_3d()
    s = w.setup(Solver(3, "musl", ordered=ordered))
    s.step(STEPS, w.dt)
    p = s.particles(["x", "v", "stress", "cell", "xi"])
    s.close()
    return w, p


def test_wedge_full_size_binning_and_fast_vs_ordered():
    w, fast = _run(False)
    assert len(fast["x"]) == 16806912
    h, n = 0.02, np.array([256, 128, 128])
    ijk = np.floor(fast["x"] / h).astype(np.int64)
    assert np.all(ijk >= 0) and np.all(ijk < n)
    cell = ijk[:, 0] + n[0] * (ijk[:, 1] + n[1] * ijk[:, 2])
    assert np.array_equal(cell, fast["cell"])
    xi = 2.0 * (fast["x"] - (ijk + 0.5) * h) / h
    assert np.max(np.abs(xi - fast["xi"])) < 1e-9
    moved = np.max(np.abs(fast["v"]))
    assert np.isfinite(moved) and moved > 0.0
    _, ordered = _run(True)
    assert np.array_equal(ordered["cell"], fast["cell"])
    for f in ("x", "v", "stress"):
        scale = np.max(np.abs(ordered[f]))
        err = np.max(np.abs(fast[f] - ordered[f])) / scale
        assert err <= 1e-12 * STEPS, (f, err)
