"""Compile the HIP library (libradeonrays.so) in-tree for gfx950.

Every source under csrc/ is compiled with hipcc for gfx950 only, with
floating-point contraction off (results are a fixed function of the inputs)
and packed-FP32 instruction selection disabled, then linked into one shared
library next to this file.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
INCLUDE = PKG_DIR.parent / "include"
LIB_PATH = PKG_DIR / "libradeonrays.so"
OBJ_DIR = PKG_DIR / "_obj"
# Host AddressSanitizer + UndefinedBehaviorSanitizer build of the same
# sources (host code only: each -fsanitize= goes to the host pass through
# -Xarch_host; device code is compiled as usual).
ASAN_LIB_PATH = PKG_DIR / "libradeonrays_asan.so"
ASAN_OBJ_DIR = PKG_DIR / "_obj_asan"
HOST_SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                 "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
                 "-Xarch_host", "-shared-libsan"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Only the rr*/rrx* entry points (declared RR_API, i.e. default visibility,
# in include/*.h) are exported; everything else stays inside the library.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    "-fvisibility-inlines-hidden",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _sources():
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


def _compile(src: Path, verbose: bool, sanitize: bool = False) -> Path:
    obj = (ASAN_OBJ_DIR if sanitize else OBJ_DIR) / (src.stem + ".o")
    cmd = [HIPCC, *DEVICE_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-fPIC", "-fvisibility=hidden",
               "-fvisibility-inlines-hidden", "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(src), "-o",
               str(obj)]
    if sanitize:
        cmd[1:1] = HOST_SANITIZE
    proc = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in proc.stderr.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{errs}")
    if verbose and errs.strip():
        print(errs, file=sys.stderr)
    return obj


def build(verbose: bool = False, jobs: int | None = None, sanitize: bool = False) -> Path:
    """Compile every source and link LIB_PATH (ASAN_LIB_PATH with sanitize);
    always rebuilds."""
    obj_dir, lib_path = (ASAN_OBJ_DIR, ASAN_LIB_PATH) if sanitize else (OBJ_DIR, LIB_PATH)
    obj_dir.mkdir(exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(len(srcs), os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose, sanitize), srcs))
    tmp = lib_path.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", f"-Wl,--version-script={CSRC / 'exports.map'}",
           *map(str, objs), "-o", str(tmp)]
    if sanitize:
        cmd[1:1] = HOST_SANITIZE
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stderr}")
    os.replace(tmp, lib_path)
    return lib_path


if __name__ == "__main__":
    print(build(verbose=True, sanitize="--sanitize" in sys.argv))
