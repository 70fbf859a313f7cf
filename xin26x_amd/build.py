"""Builds libxg266.so (every HIP source under csrc/, gfx950) in-tree.

    python -m xin26x_amd.build          # or __graft_entry__.build()

hipcc compiles for gfx950 without a GPU present.  Packed-FP32 instruction
selection is disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import concurrent.futures
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libxg266.so")
OBJDIR = os.path.join(ROOT, "_build", "xg266")
ARCH = os.environ.get("XG_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-amdgpu-mfma-vgpr-form",      # MFMA results in VGPRs: no accumulator-register reads
         "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include")]


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


def build_oracle(verbose=True):
    """The CPU checker (oracle/liboracle.so) and, when the xin26x sources are
    present, oracle/_ref/libxin_ref.so (10-bit C) and libxin_ref8.so (8-bit,
    with the SSSE3/AVX2 table entries).  Test infrastructure only."""
    od = os.path.join(ROOT, "oracle")
    subprocess.check_call(["make", "-s", "-C", od, "-B", "liboracle.so"])
    ref = os.environ.get("XG_REFERENCE_DIR", "/root/reference")
    if os.path.isdir(os.path.join(ref, "h266_core")):
        subprocess.check_call(["make", "-s", "-C", od, "-B", "ref", f"REF={ref}"])
    if verbose:
        print("built oracle")


if __name__ == "__main__":
    build()
    if "--oracle" in sys.argv:
        build_oracle()
