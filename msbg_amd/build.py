"""Build step of libmsbg_gpu.so (HIP, gfx950) and the CPU oracle.

Compiles every source under msbg_amd/csrc with hipcc into object files in
msbg_amd/_build (in parallel), links them into msbg_amd/libmsbg_gpu.so, and
builds oracle/ with make. Run as `python -m msbg_amd.build`.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "msbg_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmsbg_gpu.so")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 (v_pk_*_f32) stays off: the hardware path it selects has
    # given silent wrong results on these boxes
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT}/include",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *HIPFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_lib(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle(verbose=True):
    odir = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-B", "-C", odir], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {odir}/liboracle_strict.so, liboracle_fast.so, libcnn_oracle.so, libps_oracle.so")


def build_facade_demo(verbose=True):
    """The C++ facade example used by the tests (host code only)."""
    src = os.path.join(ROOT, "tests", "cpp", "facade_demo.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "facade_demo")
    cmd = ["g++", "-O2", "-std=c++17", "-o", out, src, f"-L{PKG}", "-lmsbg_gpu",
           f"-Wl,-rpath,$ORIGIN/../../msbg_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"facade demo build failed: {' '.join(cmd)}\n{r.stderr}")
    if verbose:
        print(f"built {out}")


def main():
    build_oracle()
    build_lib()
    build_facade_demo()


if __name__ == "__main__":
    sys.exit(main())
