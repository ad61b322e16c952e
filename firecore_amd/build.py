"""Build libMMFFmulti_lib.so (HIP kernels for gfx950 + host C++ + C-ABI) in-tree.

    python -m firecore_amd.build        # or firecore_amd.build.build()
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libMMFFmulti_lib.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# packed FP32 instructions are disabled for all device code
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wno-unused-command-line-argument",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HIP_FLAGS = ["-x", "hip", "--offload-arch=" + ARCH, "-munsafe-fp-atomics", "-fno-hip-fp32-correctly-rounded-divide-sqrt"]
HOST_FLAGS = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]

HIP_SOURCES = ["relax_persistent.hip", "split_kernels.hip", "sample_kernels.hip", "proposal_train.hip",
               "gridff_gen.hip"]
CPP_SOURCES = ["molgen.cpp", "topology.cpp", "gridff_host.cpp", "builder.cpp", "database.cpp", "world.cpp", "train.cpp", "capi.cpp"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + HIP_FLAGS + ["-c", path, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + HOST_FLAGS + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = HIP_SOURCES + CPP_SOURCES
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
