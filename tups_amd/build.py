"""Build the gfx950 shared library ``tups_amd/libtups_gpu.so`` with hipcc.

Every source is compiled on each call (this is also the build check: hipcc
cross-compiles for gfx950 without a GPU).  Objects go to ``tups_amd/_build``;
the library is linked in-tree so that it travels with the repository.

    python -m tups_amd.build            # build
    python -m tups_amd.build --fortran  # also compile the ISO_C_BINDING modules
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libtups_gpu.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["context.cpp", "tables.cpp", "iteration.cpp", "prms.cpp", "bse.hip", "phir.hip", "eqmatrix.hip",
           "crossing.hip", "vertex_ops.hip", "bse_split.hip", "bse_closed.hip", "chi.hip", "sde.hip", "gtau.hip",
           "conv.hip", "sus.hip", "uneven.hip"]

# Packed FP32 instructions are disabled for every device compile.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-I", str(ROOT / "include"),
          "-Wall", "-Wno-unused-function", "-Wno-unused-result"]


def _compile(src: str) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


HDF5_ROOT = Path("/opt/conda")
H5_LIB = PKG / "libtups_h5.so"


def build_h5(verbose: bool = False):
    """libtups_h5.so: the HDF5 vertex dump (host C++ over the HDF5 C library and
    libtups_gpu.so).  Skipped, with a message, where the HDF5 headers are absent."""
    if not (HDF5_ROOT / "include" / "hdf5.h").exists() or not (HDF5_ROOT / "lib" / "libhdf5.so").exists():
        if verbose:
            print(f"HDF5 not found under {HDF5_ROOT}: libtups_h5.so not built")
        return None
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}",
           f"-I{HDF5_ROOT / 'include'}", str(CSRC / "vertex_h5.cpp"), f"-L{PKG}", "-ltups_gpu",
           f"-L{HDF5_ROOT / 'lib'}", "-lhdf5", f"-Wl,-rpath,$ORIGIN:{HDF5_ROOT / 'lib'}", "-o", str(H5_LIB)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"libtups_h5.so build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {H5_LIB}")
    return H5_LIB


def build_fortran(verbose: bool = False) -> Path:
    """Compile the Fortran ISO_C_BINDING interface + drop-in modules, the one-iteration driver
    and the whole-run program tups_run."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    fdir = PKG / "fortran"
    out = BUILD / "fortran"
    out.mkdir(parents=True, exist_ok=True)
    srcs = ["tups_gpu_iface.f90", "tups_state.f90", "tups_dropin.f90", "tups_driver.f90"]
    objs = []
    for s in srcs:
        obj = out / (Path(s).stem + ".o")
        cmd = [fc, "-O2", "-J", str(out), "-c", str(fdir / s), "-o", str(obj)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed for {s}:\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    exe = PKG / "tups_fortran_driver"
    cmd = [fc, *map(str, objs), f"-L{PKG}", "-ltups_gpu", f"-Wl,-rpath,$ORIGIN", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran link failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {exe}")
    # the whole run from a parameter file
    obj = out / "tups_run.o"
    r = subprocess.run([fc, "-O2", "-J", str(out), "-c", str(fdir / "tups_run.f90"), "-o", str(obj)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed for tups_run.f90:\n{r.stdout}\n{r.stderr}")
    run = PKG / "tups_run"
    r = subprocess.run([fc, str(out / "tups_gpu_iface.o"), str(obj), f"-L{PKG}", "-ltups_gpu", "-Wl,-rpath,$ORIGIN",
                        "-o", str(run)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran link failed for tups_run:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {run}")
    return exe


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--fortran", action="store_true")
    args = ap.parse_args(argv)
    build(verbose=True)
    if args.fortran:
        build_fortran(verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
