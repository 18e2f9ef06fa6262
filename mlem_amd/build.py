"""Build recipe for the native parts of the MLEM engine.

Products (all in-tree, next to the sources that need them):
  mlem_amd/lib/libmlem.so        HIP kernels for gfx950 + the C ABI (include/mlem.h)
  mlem_amd/lib/libmlem_synth.so  synthetic matrices / phantoms / sinograms (host)
  mlem_amd/bin/mlem_cli          drop-in command line driver

Run `python -m mlem_amd.build` or call build_all().
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "mlem_amd"
CSRC = PKG / "csrc"
LIB = PKG / "lib"
BIN = PKG / "bin"
OBJ = PKG / "lib" / "obj"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Flags shared by every HIP translation unit:
#  -ffp-contract=off       multiply/add fuse only where a kernel asks for fmaf
#  -fdenormal-fp-math-f32=preserve-sign,ieee   fp32 subnormal results flushed,
#                          subnormal inputs kept (the reference sets FTZ, not DAZ)
#  -packed-fp32-ops off    no packed fp32 instructions are generated
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-ffp-contract=off", "-Xclang", "-fdenormal-fp-math-f32=preserve-sign,ieee",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{CSRC}"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "is not a recognized feature" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def _hip_object(src: Path) -> Path:
    obj = OBJ / (src.name + ".o")
    lang = [] if src.suffix == ".hip" else ["-x", "hip"]
    _run([HIPCC, *HIP_FLAGS, *lang, "-c", str(src), "-o", str(obj)])
    return obj


def build_libmlem() -> Path:
    srcs = [CSRC / "kernels.hip", CSRC / "layout.hip", CSRC / "batch.hip", CSRC / "unroll.hip",
            CSRC / "api.cpp", CSRC / "csr4.cpp", CSRC / "symmetry.cpp", CSRC / "stream.cpp",
            CSRC / "panel.hip", CSRC / "panels.cpp", CSRC / "noalloc.cpp"]
    with ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        objs = list(ex.map(_hip_object, srcs))
    out = LIB / "libmlem.so"
    tmp = LIB / "libmlem.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-Wl,-z,defs"])
    os.replace(tmp, out)
    return out


def build_synth() -> Path:
    out = LIB / "libmlem_synth.so"
    tmp = LIB / "libmlem_synth.so.tmp"
    _run(["g++", *CXX_FLAGS, "-fopenmp", "-shared", "-o", str(tmp),
          str(CSRC / "synth.cpp"), str(CSRC / "csr4.cpp")])
    os.replace(tmp, out)
    return out


def build_cli() -> Path:
    out = BIN / "mlem_cli"
    tmp = BIN / "mlem_cli.tmp"
    _run(["g++", *CXX_FLAGS, "-o", str(tmp), str(CSRC / "cli.cpp"), str(CSRC / "csr4.cpp"),
          f"-L{LIB}", "-lmlem", "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, out)
    return out


def build_all(verbose: bool = True) -> None:
    for d in (LIB, BIN, OBJ):
        d.mkdir(parents=True, exist_ok=True)
    with ThreadPoolExecutor(max_workers=2) as ex:
        f_lib = ex.submit(build_libmlem)
        f_syn = ex.submit(build_synth)
        for f in (f_lib, f_syn):
            p = f.result()
            if verbose:
                print(f"built {p.relative_to(ROOT)}")
    p = build_cli()
    if verbose:
        print(f"built {p.relative_to(ROOT)}")
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    try:
        build_all()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
