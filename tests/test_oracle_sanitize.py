"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the
driver oracle/sanitize_check.c, compiled together with oracle/sph_oracle.c,
runs BGGS and leap-frog steps through every switch of the oracle on small
seeded sets in 2D and 3D, and must finish with no sanitizer report. A build
with a planted read past the end of an array shows that the sanitizer is
live (it must stop that run)."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "sanitize_check.c")
FLAGS = ["-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-ffp-contract=off", "-Wall", "-Wno-unknown-pragmas"]

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")


def _build(tmp_path, *defs):
    exe = tmp_path / ("sancheck" + "".join(d.replace("-D", "_") for d in defs))
    r = subprocess.run(["gcc", *FLAGS, *defs, "-I" + os.path.join(ROOT, "oracle"), SRC, "-o", str(exe),
                        "-lm"], capture_output=True, text=True)
    if r.returncode != 0 and "asan" in (r.stderr + r.stdout).lower():
        pytest.skip("the sanitizer runtime is not installed")
    assert r.returncode == 0, r.stderr
    return exe


def _run(exe):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    return subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)


def test_oracle_is_clean_under_address_and_undefined_sanitizers(tmp_path):
    r = _run(_build(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert "sanitized oracle: all cases ok" in r.stdout


def test_planted_overflow_is_caught(tmp_path):
    r = _run(_build(tmp_path, "-DSANITIZE_PLANT"))
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stderr
