"""The oracle's CPU tests again, against its AddressSanitizer + UndefinedBehaviorSanitizer
build (`make -C oracle sanitize`: host code only, python with libasan preloaded).  Any
ASan report or UBSan error aborts that run, so this test fails on it."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _libasan():
    cc = shutil.which("gcc")
    if not cc:
        return None
    p = subprocess.run([cc, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout
    p = p.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.mark.skipif(_libasan() is None, reason="gcc's libasan is not installed")
def test_oracle_tests_pass_under_asan_and_ubsan():
    env = dict(os.environ)
    env.pop("NBX_ORACLE_SO", None)
    res = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                         capture_output=True, text=True, env=env, timeout=600)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert " passed" in out
