"""In-tree build of libndg_amd.so for gfx950 (hipcc), used by __graft_entry__.build().

Kernels (*.hip) are compiled with hipcc's default FP contraction; the host sources (*.cpp), which
reproduce the reference host arithmetic (tables, face frames, time accumulation), are compiled
with -ffp-contract=off. Packed FP32 instructions are disabled for all device code.
"""
import concurrent.futures
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libndg_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-std=c++17", "-fPIC", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
DEVICE_FLAGS = ["--offload-arch=" + ARCH, "-O3"] + COMMON
HOST_FLAGS = ["-O2", "-ffp-contract=off", "-Wall"] + COMMON


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 target feature and says so; that is harmless
    noise = [l for l in (proc.stdout + proc.stderr).splitlines()
             if l.strip() and "packed-fp32-ops" not in l and "warning generated" not in l
             and "warnings generated" not in l]
    if proc.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + "\n".join(noise))
    return noise


def build(verbose=False):
    """Compiles every source under csrc/ and links ndg_cuda_amd/libndg_amd.so. Always rebuilds."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for name in sorted(os.listdir(CSRC)):
        src = os.path.join(CSRC, name)
        stem, ext = os.path.splitext(name)
        if ext not in (".hip", ".cpp"):
            continue
        obj = os.path.join(BUILD, stem + ".o")
        flags = DEVICE_FLAGS if ext == ".hip" else HOST_FLAGS
        jobs.append(([HIPCC] + flags + ["-c", src, "-o", obj], obj))
    # one compiler process per source, at most as many as there are cores (16 at most)
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 16))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
        results = list(pool.map(lambda job: _run(job[0]), jobs))
    if verbose:
        for msgs in results:
            if msgs:
                print("\n".join(msgs), file=sys.stderr)
    objs = [obj for _, obj in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_host_driver()
    return LIB


HOST_DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "host_api_main.cpp")
HOST_DRIVER = os.path.join(ROOT, "tests", "cpp", "host_api_main")
FLOW_DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "reference_flow_main.cpp")
FLOW_DRIVER = os.path.join(ROOT, "tests", "cpp", "reference_flow_main")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build_host_driver():
    """Links the C++ host-API drivers (tests/cpp/host_api_main.cpp, and reference_flow_main.cpp,
    which also calls the HIP runtime for its stream) against libndg_amd.so with g++."""
    if not os.path.exists(HOST_DRIVER_SRC):
        return None
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), HOST_DRIVER_SRC,
          "-L" + PKG, "-lndg_amd", "-Wl,-rpath,$ORIGIN/../../ndg_cuda_amd", "-o", HOST_DRIVER])
    if os.path.exists(FLOW_DRIVER_SRC):
        _run(["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROOT, "include"),
              "-I" + os.path.join(ROCM, "include"), FLOW_DRIVER_SRC, "-L" + PKG, "-lndg_amd",
              "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../ndg_cuda_amd",
              "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-o", FLOW_DRIVER])
    return HOST_DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
