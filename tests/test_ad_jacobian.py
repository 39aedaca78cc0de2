"""The forward-mode Mohr-Coulomb update used by the adjoint
(mpm_amd/csrc/mpmx_ad.h), compiled for the host with g++: with T = double it
reproduces the oracle's return mapping bit for bit on random trial states
(elastic and plastic, shear and tensile, with and without cohesion,
dilation and tension cutoff), and its Dual tangents with respect to stress,
strain increment, Young's modulus and friction angle agree with central
differences (to the differences' own accuracy, 1e-4 of the largest tangent)."""
import os
import shutil
import subprocess

import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "mpm_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_mc_update_matches_oracle_and_differences(tmp_path):
    oracle.build()
    exe = str(tmp_path / "mc_jacobian_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                    "-I", CSRC, os.path.join(HERE, "ad", "mc_jacobian_check.cpp"), "-o", exe,
                    "-ldl"], check=True)
    lib = os.path.join(HERE, "..", "oracle", "liboracle.so")
    out = subprocess.run([exe, lib], check=True, capture_output=True, text=True).stdout.split()
    cases, plastic, mismatches, maxrel = int(out[0]), int(out[1]), int(out[2]), float(out[3])
    assert cases == 4000 and plastic > 1000
    assert mismatches == 0
    assert maxrel < 1e-4
