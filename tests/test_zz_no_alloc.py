"""mpmx_step_noalloc (the step with its scratch in the library's fixed
per-GPU array, no device allocation) against mpmx_step, through the C ABI:
ordered mode bit for bit, fast mode to the atomic-order rounding, on two
decks in a row (the second smaller), and its two refusals."""
import numpy as np
import pytest

from mpm_amd import workloads
from mpm_amd import materials as mat

pytestmark = pytest.mark.gpu

FIELDS = ("x", "u", "v", "stress", "strain", "F", "volume", "density", "xi", "statevars",
          "cell")


def _relerr(a, b):
    scale = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / scale if scale > 0 else float(np.max(np.abs(a)))


def _run_pair(w, ordered, nsteps):
    from mpm_amd.capi import Solver
    a = w.setup(Solver(w.dim, w.scheme, ordered=ordered))
    b = w.setup(Solver(w.dim, w.scheme, ordered=ordered))
    a.step_noalloc(nsteps, w.dt)
    b.step(nsteps, w.dt)
    assert a.step_count == b.step_count == nsteps
    pa, pb = a.particles(), b.particles()
    return a, b, pa, pb


def _decks():
    big = workloads.collapse_3d(nx=12, ny=6, nz=6, block=(6, 6, 6))   # 1728 particles
    small = workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))  # 512 particles
    return big, small


def test_noalloc_equals_step_ordered_twice():
    big, small = _decks()
    for w in (big, small):
        print("noalloc ordered", w.name, len(w.cell_set) * 8, flush=True)
        a, b, pa, pb = _run_pair(w, True, 40)
        for f in FIELDS:
            assert np.array_equal(pa[f], pb[f]), f
        # the context keeps stepping with either entry point afterwards
        a.step(3, w.dt)
        b.step_noalloc(3, w.dt)
        pa, pb = a.particles(), b.particles()
        for f in FIELDS:
            assert np.array_equal(pa[f], pb[f]), f


def test_noalloc_equals_step_fast_twice():
    """Fast mode sums in atomic order, so two runs of either entry point
    differ by rounding that the Mohr-Coulomb steps amplify (a few 1e-10
    after 20 steps): two steps per call (a full sort, then a key refresh)."""
    big, small = _decks()
    for w in (big, small):
        print("noalloc fast", w.name, len(w.cell_set) * 8, flush=True)
        a, b, pa, pb = _run_pair(w, False, 2)
        assert np.array_equal(pa["cell"], pb["cell"])
        for f in ("x", "v", "stress", "strain", "F"):
            assert _relerr(pa[f], pb[f]) <= 1e-11, f


def test_noalloc_2d_linear_elastic_jaumann_usl():
    w = workloads.column_2d(nx=8, ny=16)
    w.materials = mat.linear_elastic(1e6, 0.3, 1800.0, stress_rate="jaumann")[None, :]
    w.scheme = "usl"
    print("noalloc 2d jaumann usl", flush=True)
    a, b, pa, pb = _run_pair(w, True, 30)
    for f in FIELDS:
        if f != "statevars":
            assert np.array_equal(pa[f], pb[f]), f


def test_noalloc_refuses_remove_mode_and_oversize():
    from mpm_amd.capi import Solver, MpmxError, E_UNSUPPORTED, E_ARG
    w = workloads.column_2d(nx=8, ny=16)
    s = w.setup(Solver(2, w.scheme))
    s.set_locate_particles(False)
    with pytest.raises(MpmxError) as e:
        s.step_noalloc(1, w.dt)
    assert e.value.code == E_UNSUPPORTED
    # 360 x 360 quad4 cells, 4 Mohr-Coulomb particles each: 518,400 particles
    # need ~71 MiB of scratch, more than the 64 MiB array
    from mpm_amd.mesh import structured_mesh
    coords, cells = structured_mesh((0.0, 0.0), (0.01, 0.01), (360, 360))
    s = Solver(2, "usl")
    s.set_mesh(coords, cells)
    s.set_materials(mat.mohr_coulomb(1e6, 0.3, 1800.0, friction=30.0, dilation=0.0,
                                     cohesion=1e3)[None, :])
    s.generate_gauss(2)
    s.compute_volume_mass()
    print("noalloc oversize", s.nparticles, flush=True)
    with pytest.raises(MpmxError) as e:
        s.step_noalloc(1, 1e-5)
    assert e.value.code == E_ARG and "scratch" in str(e.value)
    s.step(1, 1e-5)  # the allocating entry point still runs it
