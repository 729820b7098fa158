"""Build recipe of libfvs_hip.so (HIP, gfx950) and of the CPU oracle.

The library is compiled in-tree with hipcc; no JIT cache is involved, so the
built files travel with the source tree.  FP64 parity with the reference
requires that no multiply-add is fused (-ffp-contract=off) and that FP64
division and square root stay IEEE (no fast-math flags); only the
translation units of the tolerance-checked arithmetic modes fuse multiply-adds.
Packed FP32 VALU instructions are disabled (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libfvs_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["fvs_api.hip", "fvs_sweep.hip", "fvs_sweep_fast.hip", "fvs_godunov.hip", "fvs_mlp.hip",
           "fvs_kbn.hip", "fvs_batch.hip", "fvs_batch_fast.hip", "fvs_amr.hip",
           "fvs_level_composed.hip"]

# Translation units that hold only the kernels of the non-IEEE arithmetic
# modes (FVS_FP64_RCP, FVS_FP32_FLUX): there multiply-adds may be fused,
# since those modes are held to a tolerance, not to bit equality.
CONTRACTED = {"fvs_sweep_fast.hip", "fvs_batch_fast.hip"}

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _objects(jobs: int = 8) -> list[Path]:
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        objs.append(obj)
        extra = ["-ffp-contract=fast"] if src in CONTRACTED else []
        cmd = [HIPCC, *HIP_FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError(f"hipcc failed on {src}")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    return objs


def build_library() -> Path:
    LIBDIR.mkdir(exist_ok=True)
    objs = _objects()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


HDF5_PREFIX = Path("/opt/conda")
HDF5_LIB = LIBDIR / "libfvs_hdf5.so"


def build_hdf5_writer() -> Path | None:
    """Host-only HDF5 output library (io/fvs_hdf5.cpp) against the HDF5 C
    library of the image; skipped when its headers are absent."""
    if not (HDF5_PREFIX / "include" / "hdf5.h").exists():
        print("[build] no HDF5 headers: libfvs_hdf5.so not built", flush=True)
        return None
    LIBDIR.mkdir(exist_ok=True)
    tmp = HDF5_LIB.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{ROOT / 'include'}",
          f"-I{HDF5_PREFIX / 'include'}", str(PKG / "io" / "fvs_hdf5.cpp"),
          f"-L{HDF5_PREFIX / 'lib'}", f"-Wl,-rpath,{HDF5_PREFIX / 'lib'}", "-lhdf5",
          "-o", str(tmp)])
    os.replace(tmp, HDF5_LIB)
    return HDF5_LIB


def build_oracle() -> Path:
    """Compiles the CPU oracle (test infrastructure) with its own recipe."""
    _run(["make", "-C", str(ROOT / "oracle"), "-j4"])
    return ROOT / "oracle" / "libfvs_oracle.so"


def main(argv: list[str]) -> int:
    what = argv[1:] or ["lib", "oracle"]
    if "lib" in what:
        build_library()
    if "oracle" in what:
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
