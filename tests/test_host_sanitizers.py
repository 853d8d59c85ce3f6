"""The C-ABI layer and the oracle under host AddressSanitizer and
UndefinedBehaviorSanitizer (the options the reference's
cmake/Sanitizers.cmake:1-48 offers for its host code).

Both libraries are rebuilt with the sanitizers on their host code
(libradeonrays_asan.so: every -fsanitize= passed to hipcc through
-Xarch_host, device code unchanged; liboracle_asan.so: ROCm's clang++, so
both use one shared sanitizer runtime), then the CPU
suites that exercise them — tests/test_abi.py (every entry point's argument
checks, handle registry, memory requirements) and tests/test_oracle.py (the
oracle against its closed forms, brute force and the reference's
bvh_analyzer) — run in a child interpreter with the sanitizer runtime
preloaded.  Any report (heap overflow, use after free, undefined behaviour)
aborts the child and fails this test.  Leak checking is off: the Python
interpreter itself does not free everything at exit.
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _runtime() -> str:
    """ROCm clang's shared sanitizer runtime (ASan with the UBSan handlers),
    which both sanitized libraries link against."""
    out = subprocess.run(["/opt/rocm/llvm/bin/clang++", "-print-file-name=libclang_rt.asan-x86_64.so"],
                         capture_output=True, text=True).stdout.strip()
    if not out or not os.path.isabs(out) or not Path(out).exists():
        hits = sorted(Path("/opt/rocm/llvm/lib/clang").glob("*/lib/linux/libclang_rt.asan-x86_64.so"))
        if not hits:
            pytest.skip("libclang_rt.asan-x86_64.so not found")
        out = str(hits[-1])
    return out


def test_abi_and_oracle_under_asan_and_ubsan():
    sys.path.insert(0, str(ROOT))
    from radeonrays_sdk_amd import _build

    lib = _build.build(sanitize=True)
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "liboracle_asan.so"], check=True)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": _runtime(),
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1:halt_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "RADEONRAYS_LIB": str(lib),
        "RR_ORACLE_LIB": str(ROOT / "oracle" / "liboracle_asan.so"),
        "OMP_NUM_THREADS": "4",
    })
    proc = subprocess.run(
        [sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
         "tests/test_abi.py", "tests/test_oracle.py"],
        cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (proc.stdout + proc.stderr)[-4000:]
    assert proc.returncode == 0, tail
    assert "passed" in proc.stdout and "AddressSanitizer" not in proc.stderr and "runtime error" not in proc.stderr, tail
