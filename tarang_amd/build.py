"""Build the gfx950 shared library libtarang_amd.so in-tree.

Compiles every HIP source under tarang_amd/csrc with hipcc for gfx950 and
links them into tarang_amd/libtarang_amd.so, whose exported symbols are the
C ABI declared in include/tarang_gpu.h; builds the HDF5 field-dump library
tarang_amd/libtarang_io.so (include/tarang_io.h) and the C++ drivers.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtarang_amd.so")
DRIVER = os.path.join(HERE, "tarang_ifluid")
DRIVER_RBC = os.path.join(HERE, "tarang_rbc")
IO_LIB = os.path.join(HERE, "libtarang_io.so")
CXX = os.environ.get("CXX", "g++")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HDF5 = os.environ.get("TG_HDF5_ROOT", "/opt/conda")
SOURCES = ["tg_kernels.hip", "tg_diag.hip", "tg_mhd.hip", "tg_adjoint.hip", "tg_sff.hip", "tg_chff.hip", "tg_api.hip"]
# Sources whose kernels are templated on the transform length: compiled once
# more per length unit (TG_LEN_UNIT 1: power-of-two FP64, 2: power-of-two
# FP32, 3 and 4: radix-3/5 lengths FP64), so that the instances compile in
# parallel; the plain compilation holds their launchers.
LENGTH_SOURCES = ["tg_kernels.hip", "tg_mhd.hip", "tg_adjoint.hip", "tg_sff.hip"]
LENGTH_UNITS = (1, 2, 3, 4)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TG_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _compile(job):
    src, unit = job
    stem = os.path.splitext(src)[0] + (f"_u{unit}" if unit else "")
    out = os.path.join(OBJ, stem + ".o")
    cmd = [HIPCC] + FLAGS + ([f"-DTG_LEN_UNIT={unit}"] if unit else []) + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = [(src, u) for src in LENGTH_SOURCES for u in LENGTH_UNITS] + [(src, 0) for src in SOURCES]
    # largest first; one hipcc per host core (each takes 1-3 GB)
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, min(len(jobs), os.cpu_count() or 8))) as ex:
        objs = list(ex.map(_compile, jobs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # C++ host drivers over the C ABI (tarang_amd.hpp)
    for src, out in (("ifluid_main.cpp", DRIVER), ("rbc_main.cpp", DRIVER_RBC)):
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
               os.path.join(HERE, "host", src), "-o", out + ".tmp", "-L", HERE, "-ltarang_amd", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed for {src}:\n{r.stdout}\n{r.stderr}")
        os.replace(out + ".tmp", out)
    build_io()
    if verbose:
        print(f"built {LIB}, {IO_LIB}, {DRIVER} and {DRIVER_RBC}")
    return LIB


def build_io():
    """libtarang_io.so: the HDF5 field dumps (host C++ over the HDF5 C library
    of the image's /opt/conda and the HIP runtime for device copies)."""
    cmd = [CXX, "-O2", "-fPIC", "-shared", "-std=c++17", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROCM, "include"), "-I", os.path.join(HDF5, "include"),
           os.path.join(HERE, "io", "tg_io.cpp"), "-o", IO_LIB + ".tmp",
           "-L", os.path.join(HDF5, "lib"), "-lhdf5", "-Wl,-rpath," + os.path.join(HDF5, "lib"),
           "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"io library build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(IO_LIB + ".tmp", IO_LIB)
    return IO_LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
