"""Build libprodm_mdr.so (gfx950) in-tree with hipcc.

    python -m prodm_amd.build [--jobs N]

Every call recompiles. Flags: exact IEEE arithmetic as written (no FMA
contraction, no denormal flushing) so that the element-wise kernels round
like the CPU restatement, and no packed-FP32 instructions.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libprodm_mdr.so")
SOURCES = ["capi.hip", "mgard.hip", "bitplane.hip", "interleave.hip", "collect.hip", "adjoint.hip", "batched.hip",
           "noalloc.hip"]
HOST_SOURCES = ["planner.cpp", "perbit_host.cpp"]  # host-only C++ (g++)
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cpp"):
        cmd = [os.environ.get("CXX", "g++"), *HOST_FLAGS, f"-I{INCLUDE}", "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [HIPCC, *FLAGS, f"-I{INCLUDE}", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj, r.stderr


ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ZSTD_INC = os.environ.get("ZSTD_INC", "/opt/conda/include")  # zstd.h (libzstd.so.1 is the system library)
BIN = os.path.join(HERE, "bin")
HOST_LIB = os.path.join(HERE, "libprodm_host.so")
# host C++ programs over the library: (output, source)
PROGRAMS = [
    ("prodm_refactor", os.path.join(HERE, "apps", "prodm_refactor.cpp")),
    ("prodm_reconstruct", os.path.join(HERE, "apps", "prodm_reconstruct.cpp")),
    ("test_cpp_pipeline", os.path.join(os.path.dirname(HERE), "tests", "cpp", "test_cpp_pipeline.cpp")),
]


def _cxx_host(out, src, shared=False, link_core=True):
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{INCLUDE}", f"-I{ROCM}/include",
           "-idirafter", ZSTD_INC, src, "-o", out]
    if shared:
        cmd[1:1] = ["-shared", "-fPIC"]
    if link_core:
        cmd += [f"-L{HERE}", "-lprodm_mdr", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/.."]
    cmd += [f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib", "-l:libzstd.so.1", "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{cxx} failed for {os.path.basename(src)}:\n{r.stderr}")
    return out


def build_host(jobs=None):
    """Host C++ over the library: libprodm_host.so, the apps and the C++ tests."""
    os.makedirs(BIN, exist_ok=True)
    jobs = jobs or min(4, os.cpu_count() or 1)
    tasks = [(HOST_LIB + ".tmp", os.path.join(CSRC, "host_capi.cpp"), True, False)]
    tasks += [(os.path.join(BIN, name), src, False, True) for name, src in PROGRAMS]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda t: _cxx_host(*t), tasks))
    os.replace(HOST_LIB + ".tmp", HOST_LIB)


SAN_DIR = os.path.join(BUILD, "san")


def build_host_sanitized():
    """libprodm_host.so under AddressSanitizer + UndefinedBehaviorSanitizer
    (host code, g++) in prodm_amd/_build/san/, for a CPU test run with
    PRODM_HOST_LIB pointing at it and gcc's libasan preloaded."""
    os.makedirs(SAN_DIR, exist_ok=True)
    out = os.path.join(SAN_DIR, "libprodm_host.so")
    cmd = [os.environ.get("CXX", "g++"), "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-std=c++17", "-Wall", "-shared", "-fPIC", f"-I{INCLUDE}",
           "-idirafter", ZSTD_INC, os.path.join(CSRC, "host_capi.cpp"), "-o", out, "-l:libzstd.so.1", "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"sanitizer build failed:\n{r.stderr}")
    return out


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = SOURCES + HOST_SOURCES
    jobs = jobs or min(len(srcs), os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, srcs))
    for obj, err in results:
        if verbose and err.strip():
            sys.stderr.write(err)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_host(jobs)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--sanitize", action="store_true", help="only the sanitizer build of the host library")
    a = ap.parse_args()
    print(build_host_sanitized() if a.sanitize else build(a.jobs, a.verbose))
