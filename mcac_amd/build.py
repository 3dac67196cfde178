"""Builds the MI355X libraries in-tree.

  mcac_amd/lib/libmcac_gpu.so   HIP kernels + C-ABI (include/mcac_gpu.h), gfx950
  mcac_amd/lib/libmcac_sim.so   host driver (include/mcac_sim.h), g++, links libmcac_gpu
  mcac_amd/bin/mcac_run         command-line driver: mcac_run <params.ini> [output_dir]

FP policy on both sides: no contraction, no fast-math (IEEE FP64 as written).
"""
import os
import pathlib
import shutil
import subprocess

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "lib"
BIN = PKG / "bin"
CSRC = PKG / "csrc"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else "hipcc")
ARCH = os.environ.get("MCAC_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-std=c++17", "-O3", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-Wall", "-Wextra"]
# HDF5 C library for the snapshot files (present in the image under /opt/conda)
HDF5_ROOT = pathlib.Path(os.environ.get("MCAC_HDF5_ROOT", "/opt/conda"))


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False) -> None:
    LIB.mkdir(exist_ok=True)
    BIN.mkdir(exist_ok=True)
    gpu = LIB / "libmcac_gpu.so"
    tmp = LIB / "libmcac_gpu.so.tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", str(tmp), str(CSRC / "mcac_gpu.hip")])
    os.replace(tmp, gpu)
    host_src = [str(CSRC / "host" / f) for f in ("physics.cpp", "pick_table.cpp", "sim.cpp", "batch.cpp",
                                                  "xdmf_writer.cpp")]
    if not (HDF5_ROOT / "include" / "hdf5.h").exists():
        raise RuntimeError(f"HDF5 headers not found under {HDF5_ROOT} (set MCAC_HDF5_ROOT)")
    sim = LIB / "libmcac_sim.so"
    tmp = LIB / "libmcac_sim.so.tmp"
    _run([CXX, *CXX_FLAGS, f"-I{HDF5_ROOT / 'include'}", "-shared", "-o", str(tmp), *host_src, f"-L{LIB}",
          "-lmcac_gpu", f"-L{HDF5_ROOT / 'lib'}", "-lhdf5", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{HDF5_ROOT / 'lib'}"])
    os.replace(tmp, sim)
    exe = BIN / "mcac_run"
    _run([CXX, *CXX_FLAGS, "-o", str(exe), str(CSRC / "host" / "main.cpp"), f"-L{LIB}", "-lmcac_sim",
          "-lmcac_gpu", "-Wl,-rpath,$ORIGIN/../lib"])
    if verbose:
        print(f"built {gpu}, {sim}, {exe}")


if __name__ == "__main__":
    build(verbose=True)
