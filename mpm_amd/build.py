"""Build recipe for the HIP library (gfx950) and the C++ host driver.

``build()`` compiles every source under mpm_amd/csrc into
``mpm_amd/libmpmx.so`` with hipcc, in-tree, so the library travels with the
repository snapshot. Flags:

* ``-ffp-contract=off``: no FMA contraction, so every FP64 expression rounds
  exactly as written (the reference's x86-64 build has no FMA either);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 ops;
* ``-munsafe-fp-atomics``: hardware FP64 atomic add on coarse-grained memory.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmpmx.so")
ARCH = os.environ.get("MPMX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["mpmx_host.cpp", "mpmx_kernels.hip", "mpmx_sort.hip", "mpmx_adjoint.hip"]
HOST_TOOL = "mpmx_run"  # C++ drop-in driver (mpm::MPMExplicit over the C-ABI)


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics", "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: %s" % " ".join(cmd))
    return r.stdout


def build(verbose=False, jobs=4):
    """Compile the library (always recompiles) and return its path."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") or src.endswith(".cpp") else []
        cmds.append([cc, "--offload-arch=" + ARCH] + COMMON + lang +
                    ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            sys.stdout.write(o)
    tmp = LIB + ".tmp"
    _run([cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_driver()
    return LIB


POD_LIB = os.path.join(HERE, "libmpmx_pod.so")


def hdf5_dir():
    """Prefix of the HDF5 C and high-level libraries (headers in include/,
    libhdf5 and libhdf5_hl in lib/): $MPMX_HDF5_DIR, else the image's
    /opt/conda."""
    for cand in (os.environ.get("MPMX_HDF5_DIR"), "/opt/conda", "/usr"):
        if cand and os.path.exists(os.path.join(cand, "include", "hdf5_hl.h")):
            return cand
    raise RuntimeError("HDF5 high-level headers not found (set MPMX_HDF5_DIR)")


def build_pod():
    """The particle checkpoint table library (mpm_amd/libmpmx_pod.so, host
    only, over HDF5's high-level table API)."""
    h5 = hdf5_dir()
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, "-O2", "-std=c++17", "-Wall", "-shared", "-fPIC",
          "-I", os.path.join(HERE, "..", "include"), "-I", os.path.join(h5, "include"),
          os.path.join(HERE, "host", "mpmx_pod.cpp"), "-o", POD_LIB + ".tmp",
          "-L", os.path.join(h5, "lib"), "-lhdf5_hl", "-lhdf5",
          "-Wl,-rpath," + os.path.join(h5, "lib")])
    os.replace(POD_LIB + ".tmp", POD_LIB)
    return POD_LIB


def build_driver():
    """The C++ analysis driver (mpm_amd/mpmx_run) over the C ABI, writing and
    resuming from the particle checkpoint table."""
    build_pod()
    out = os.path.join(HERE, HOST_TOOL)
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(hipcc())))
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "..", "include"),
          os.path.join(HERE, "host", "mpmx_run.cpp"), "-o", out + ".tmp",
          "-L", HERE, "-lmpmx", "-lmpmx_pod", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath-link," + os.path.join(rocm, "lib")])
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
