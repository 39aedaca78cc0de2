"""Scatter chunk sizes (particles per workgroup of the LDS-staged scatter
kernels) and the autotuner that picks chunk and sort interval.

Every chunk size is a summation order of the same sums: from the oracle's
state, one fast-mode step with chunk 128, 256 or 512 agrees with the
oracle's step to 1e-12 (normwise per field) under USF, USL and MUSL, 2D and
3D. The autotuner leaves the particle state and step counter exactly as they
were, and a run after it agrees with a run without it to the rounding of the
changed summation order (LinearElastic column, 1e-10 over 100 steps)."""
import numpy as np
import pytest

import oracle
from mpm_amd import materials as mat
from mpm_amd import workloads
from test_gpu_parity import FIELDS, _compare, _gpu_from, _relerr

pytestmark = pytest.mark.gpu


def _collapse():
    return workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))


@pytest.mark.parametrize("chunk", [128, 512])
@pytest.mark.parametrize("dim,scheme", [(2, "usf"), (2, "usl"), (2, "musl"), (3, "musl"),
                                        (3, "usf")])
def test_chunk_per_step(chunk, dim, scheme):
    w = workloads.column_2d(nx=8, ny=16) if dim == 2 else _collapse()
    o = w.setup(oracle.Oracle(dim, scheme))
    o.step(5, w.dt)
    for s in range(5):
        g = _gpu_from(o, w, scheme)
        g.set_chunk(chunk)
        g.step(1, w.dt)
        o.step(1, w.dt)
        _compare(g, o, exact=False, what="chunk %d %s step %d" % (chunk, scheme, s))
        g.close()


def test_set_chunk_rejects_other_sizes():
    from mpm_amd.capi import MpmxError, Solver
    s = Solver(2, "usl")
    with pytest.raises(MpmxError):
        s.set_chunk(300)
    s.close()


def test_autotune_restores_state_and_picks_a_candidate():
    from mpm_amd.capi import Solver
    w = _collapse()
    a = w.setup(Solver(3, w.scheme))
    a.step(10, w.dt)
    before = a.particles()
    step0 = a.step_count
    choice = a.autotune(w.dt, trial_steps=5)
    assert choice["chunk"] in (128, 256, 512) and choice["sort_every"] in (4, 8, 16)
    after = a.particles()
    assert a.step_count == step0
    for f in FIELDS + ("cell", "statevars"):
        assert np.array_equal(before[f], after[f]), f
    a.close()


def test_run_after_autotune_matches_run_without():
    from mpm_amd.capi import Solver
    w = workloads.column_2d(nx=8, ny=16)
    a = w.setup(Solver(2, w.scheme))
    b = w.setup(Solver(2, w.scheme))
    a.autotune(w.dt, trial_steps=5)
    for _ in range(5):
        a.step(20, w.dt)
        b.step(20, w.dt)
        pa, pb = a.particles(), b.particles()
        assert np.array_equal(pa["cell"], pb["cell"])
        for f in FIELDS:
            assert _relerr(pa[f], pb[f]) <= 1e-10, f


def test_autotune_ordered_mode_is_a_no_op():
    from mpm_amd.capi import Solver
    w = workloads.column_2d(nx=8, ny=16)
    s = w.setup(Solver(2, "usf", ordered=True))
    assert s.autotune(w.dt, trial_steps=2) == dict(chunk=256, sort_every=8)
    s.close()
