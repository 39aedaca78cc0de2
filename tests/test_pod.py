"""The particle checkpoint table (include/mpmx_pod.h): the reference's
PODParticle HDF5 table (pod_particle.h:12-62, mesh.tcc:1573-1721).

CPU only: a table written from random columns reads back bit for bit, its
field names are the reference's 74 in their order (tests/golden/
pod_particle_fields.json, transcribed from src/pod_particle.cc), and the
library refuses tables with another field count.
"""
import json
import os

import numpy as np
import pytest

from mpm_amd import pod


def _random_columns(n, seed=7):
    rng = np.random.default_rng(seed)
    cols = {}
    for name, dt, k in pod.COLUMNS:
        shape = (n, k) if k > 1 else (n,)
        if dt == np.float64:
            cols[name] = rng.normal(size=shape) * 10.0 ** rng.integers(-20, 20, size=shape)
        elif dt == np.uint8:
            cols[name] = rng.integers(0, 2, size=shape).astype(np.uint8)
        elif dt == np.uint32:
            cols[name] = rng.integers(0, 2 ** 32 - 1, size=shape, dtype=np.uint64).astype(np.uint32)
        else:
            cols[name] = rng.integers(-2 ** 62, 2 ** 62, size=shape, dtype=np.int64)
    cols["nstate_vars"] = rng.integers(0, 21, size=n).astype(np.uint32)
    return cols


def test_pod_round_trip(tmp_path):
    f = str(tmp_path / "particles0100.h5")
    cols = _random_columns(2500)
    pod.write(f, cols)
    assert pod.count(f) == 2500
    back = pod.read(f)
    for name, _, _ in pod.COLUMNS:
        a, b = np.asarray(cols[name]), back[name]
        assert a.dtype == b.dtype and a.shape == b.shape, name
        if a.dtype == np.float64:  # bitwise, NaN-free by construction
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), name
        else:
            assert np.array_equal(a, b), name


def test_pod_field_names_and_order(tmp_path):
    with open(os.path.join(os.path.dirname(__file__), "golden", "pod_particle_fields.json")) as fh:
        ref = json.load(fh)
    f = str(tmp_path / "particles.h5")
    pod.write(f, _random_columns(3))
    assert pod.field_names(f) == ref["names"]
    # the record sizes follow the reference's HDF5 types
    sizes = {"llong": 8, "double": 8, "hbool": 1, "uint": 4}
    assert sum(sizes[t] for t in ref["types"]) == 8 * (1 + 39 + 9 + 1 + 20) + 2 + 8


def test_pod_empty_table_and_errors(tmp_path):
    f = str(tmp_path / "empty.h5")
    pod.write(f, {name: np.zeros((0, k) if k > 1 else 0, dtype=dt) for name, dt, k in pod.COLUMNS})
    assert pod.count(f) == 0
    with pytest.raises(RuntimeError, match="not found"):
        pod.count(str(tmp_path / "missing.h5"))
    with pytest.raises(RuntimeError, match="20"):
        bad = _random_columns(2)
        bad["nstate_vars"][:] = 21
        pod.write(str(tmp_path / "bad.h5"), bad)
