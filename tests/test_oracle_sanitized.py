"""The oracle's pins run again against its AddressSanitizer +
UndefinedBehaviorSanitizer build (host code only: `make -C oracle sanitize`),
in a child process that starts with the sanitizer runtimes preloaded: an
out-of-bounds access, a use after free or undefined behaviour in the C
restatement aborts the child and fails this test."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _runtime(name):
    try:
        out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True, check=True)
    except (OSError, subprocess.CalledProcessError):
        return None
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_pins_under_address_and_ub_sanitizers():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not (asan and ubsan):
        pytest.skip("gcc's sanitizer runtimes are not installed")
    built = subprocess.run(["make", "-C", str(ROOT / "oracle"), "-s", "sanitize"], capture_output=True, text=True)
    if built.returncode != 0:
        pytest.skip(f"the sanitizer build failed: {built.stderr[-400:]}")
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": f"{asan} {ubsan}",
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "FSI_ORACLE_LIB": str(ROOT / "oracle" / "liboracle_fsi_san.so"),
    })
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
           "tests/test_oracle_pins.py", "tests/test_phase_pins.py"]
    run = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    print(run.stdout[-2000:], run.stderr[-4000:])
    assert run.returncode == 0
    assert "passed" in run.stdout
