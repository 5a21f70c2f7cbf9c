"""Build recipe for libgps.so (gfx950) and the CPU oracle.

Compiles the HIP/C++ sources under csrc/ with hipcc for gfx950 into
gpuprefixsums_amd/libgps.so (the product library), the vendor comparator
gpuprefixsums_amd/libgps_vendor.so (rocPRIM's scan, for tests and benchmarks
only; it links libgps.so), the memory-ceiling probes
gpuprefixsums_amd/libgps_ceiling.so (benchmark only; it links libgps.so), the
reference-style test program gpuprefixsums_amd/gps_main, and the oracle (oracle/liboracle.so, test
infrastructure) with make.

    python -m gpuprefixsums_amd.build          # build everything
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"
LIB = PKG / "libgps.so"
VENDOR_LIB = PKG / "libgps_vendor.so"
CEILING_LIB = PKG / "libgps_ceiling.so"
MAIN = PKG / "gps_main"
ARCH = "gfx950"

# Library sources: kernels (.hip) and host code (.cpp, compiled as HIP so the
# runtime API and launch syntax are available).
LIB_SOURCES = [
    "scan_u32.hip",
    "segmented_u32.hip",
    "scan_fp.hip",
    "mlp.hip",
    "scan_u32x4.hip",
    "survey.hip",
    "capi.cpp",
    "capi_ext.cpp",
    "status_memory.cpp",
    "dispatchers.cpp",
    "noalloc.hip",
]

# The comparator library: rocPRIM's device scan and the CubDispatcher class.
VENDOR_SOURCES = [
    "vendor_rocprim.hip",
    "vendor_dispatcher.cpp",
]

# The memory-ceiling probes the benchmark reads the scan's roofline against.
CEILING_SOURCES = [
    "ceiling.hip",
]

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-command-line-argument",
    # Packed FP32 instructions are kept out of generated code on this target.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]

# Files with floating-point accumulation must not contract a*b+c into FMAs
# (two-sum error-free transformations rely on separately rounded ops).
PER_FILE_FLAGS = {
    "scan_fp.hip": ["-ffp-contract=off"],
    "mlp.hip": ["-ffp-contract=off"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm required to build libgps.so)")


def _compile(src: str) -> Path:
    obj = BUILD / (Path(src).name + ".o")
    cmd = [hipcc(), *COMMON_FLAGS, *PER_FILE_FLAGS.get(src, []), "-x", "hip", "-c", str(CSRC / src),
           "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def _link(objs, out: Path, extra=()) -> Path:
    tmp = out.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), *extra]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


def build_lib(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = [s for s in LIB_SOURCES + VENDOR_SOURCES + CEILING_SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(sources)))) as ex:
        objs = dict(zip(sources, ex.map(_compile, sources)))
    _link([objs[s] for s in LIB_SOURCES], LIB)
    uses_lib = [f"-L{PKG}", "-lgps", "-Wl,-rpath,$ORIGIN"]
    _link([objs[s] for s in VENDOR_SOURCES], VENDOR_LIB, uses_lib)
    _link([objs[s] for s in CEILING_SOURCES], CEILING_LIB, uses_lib)
    return LIB


def build_main() -> Path:
    src = CSRC / "gps_main.cpp"
    if not src.exists():
        return MAIN
    cmd = [hipcc(), *COMMON_FLAGS, "-x", "hip", str(src), "-o", str(MAIN), f"-L{PKG}", "-lgps", "-lgps_vendor",
           f"-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for gps_main:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return MAIN


def build_oracle() -> Path:
    res = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return ROOT / "oracle" / "liboracle.so"


def build_all() -> None:
    build_lib()
    build_main()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB}, {MAIN}, oracle", file=sys.stderr)
