"""Build recipe for librobtop_amd.so (gfx950) and the CPU oracle.

Every HIP translation unit is compiled with hipcc for gfx950 only, with packed
FP32 instructions disabled; the voxelizer and the MMA optimiser additionally
with FP contraction off (their results depend on exact FP32 rounding). Objects go to
robtop_amd/_build/, the shared library to robtop_amd/librobtop_amd.so.

    python -m robtop_amd.build          # library + oracle + tests/cpp programs
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librobtop_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]

UNITS = {
    "k_fine.hip": [],
    "k_fine_modal.hip": [],
    "k_coarse.hip": [],
    "k_galerkin.hip": [],
    "k_simp.hip": [],
    "k_worst.hip": [],
    "k_island.hip": [],
    "k_net.hip": [],
    "k_mma.hip": ["-ffp-contract=off"],
    "k_voxel.hip": ["-ffp-contract=off"],
    "rt_host.cpp": ["-x", "hip"],
    "rt_net.cpp": ["-x", "hip"],
    "rt_mma.cpp": ["-x", "hip", "-ffp-contract=off"],
}


def _compile(src, extra):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the device feature flag: drop that warning
    err = "\n".join(l for l in (p.stderr or "").splitlines() if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    return obj, err


def build_library(jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(UNITS), max(1, (os.cpu_count() or 2)))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_compile, s, e): s for s, e in UNITS.items()}
        for f in cf.as_completed(futs):
            obj, err = f.result()
            if err:
                print(f"[build] {futs[f]}: {err}", file=sys.stderr)
            objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + sorted(objs) + ["-lrocsolver", "-lrocblas"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stderr)
    os.replace(tmp, LIB)
    return LIB


def build_cpp_programs():
    """C++ programs under tests/cpp that drive the library through the C++
    facade (include/robtop_amd.hpp); each X.cpp becomes the executable X."""
    d = os.path.join(ROOT, "tests", "cpp")
    out = []
    for src in sorted(f for f in os.listdir(d) if f.endswith(".cpp")):
        exe = os.path.join(d, os.path.splitext(src)[0])
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
               os.path.join(d, src), "-L", HERE, "-lrobtop_amd", "-Wl,-rpath,$ORIGIN/../../robtop_amd", "-o", exe]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0 or p.stderr.strip():
            raise RuntimeError(f"g++ failed for {src}:\n{p.stderr}")
        out.append(exe)
    return out


def build_oracle():
    p = subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + p.stderr)
    return os.path.join(ROOT, "oracle", "liboracle.so")


if __name__ == "__main__":
    print(build_library())
    print(build_oracle())
    print("\n".join(build_cpp_programs()))
