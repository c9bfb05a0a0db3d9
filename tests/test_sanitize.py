"""The host library under AddressSanitizer and UndefinedBehaviorSanitizer.

libsufkit.so is rebuilt from the same sources with -fsanitize=address,undefined
(every report fatal) into a temporary directory, and the CPU tests that load
it run again in a child process with that build (SUFKIT_AMD_HOST_LIBRARY)
and the ASan runtime preloaded.  A sanitizer report fails the child, and so
this test.  The HIP library and the oracle are not rebuilt."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

# CPU tests of the host library (the ones that call libsufkit.so).
HOST_TESTS = ["tests/test_host_cpu.py", "tests/test_device_save.py", "tests/test_fm.py", "tests/test_prefix.py",
              "tests/test_abi.py"]


def _runtime(name: str) -> str | None:
    try:
        out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True, check=True)
    except (OSError, subprocess.CalledProcessError):
        return None
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.timeout(900)
def test_host_library_is_clean_under_asan_and_ubsan(tmp_path):
    from sufkit_amd import build

    asan = _runtime("libasan.so")
    cxx = _runtime("libstdc++.so")
    if asan is None or cxx is None or not (build.LIB / "libsufkit_hip.so").exists():
        pytest.skip("gcc's ASan runtime, libstdc++ or the HIP library is not available")
    lib = build.build_host(tmp_path / "sanitize", sanitize=True)
    env = dict(os.environ)
    env.update({
        "SUFKIT_AMD_HOST_LIBRARY": str(lib),
        # The ASan runtime first; libstdc++ right after it, so that its
        # __cxa_throw interceptor finds the real function (python does not
        # link libstdc++ itself).
        "LD_PRELOAD": f"{asan} {cxx}",
        # Python itself is not instrumented: its own allocations are not
        # leaks of the library, and it is loaded before the runtime.
        "ASAN_OPTIONS": "detect_leaks=0:verify_asan_link_order=0:abort_on_error=1",
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
    })
    proc = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                           *HOST_TESTS], cwd=ROOT, env=env, capture_output=True, text=True, timeout=850)
    tail = (proc.stdout[-3000:] + proc.stderr[-3000:])
    assert proc.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in proc.stderr and "runtime error:" not in proc.stderr, tail
    assert " passed" in proc.stdout, tail
