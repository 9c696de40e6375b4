"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
oracle/liboracle_san.so (same source, -fsanitize=address,undefined, UB
reports fatal) runs the oracle test-suite in a child process, with the
ASan runtime preloaded into the interpreter."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    cc = shutil.which("gcc")
    if not cc:
        return None
    p = subprocess.run([cc, "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_under_asan_ubsan():
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("no gcc AddressSanitizer runtime")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle_san.so"])
    env = dict(os.environ)
    env.update({"LD_PRELOAD": asan, "PQ_ORACLE_LIB": "liboracle_san.so",
                "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
                "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"})
    probe = subprocess.run([sys.executable, "-c", "from oracle import oracle as O; print(O.lib()._name)"], cwd=ROOT,
                           env=env, capture_output=True, text=True, timeout=120)
    assert probe.stdout.strip().endswith("liboracle_san.so"), probe.stdout + probe.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle.py")], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=600)
    tail = (r.stdout + r.stderr)[-3000:]
    print(tail)
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
