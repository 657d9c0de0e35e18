"""CPU: the oracle's C restatement under AddressSanitizer and
UndefinedBehaviorSanitizer.

oracle/edref.c is compiled together with tests/oracle_sanitize_driver.c by
gcc with -fsanitize=address,undefined (recovery off, so the first report
aborts) and OpenMP, and the driver calls every entry point on small sectors
of a two-orbital Hund's-coupling model with two spins and of a hybrid-bath
model.  The run must exit 0 and print OK with no sanitizer report.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "edref_san")
    cmd = ["gcc", "-std=c99", "-g", "-O1", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fopenmp",
           os.path.join(ROOT, "oracle", "edref.c"),
           os.path.join(HERE, "oracle_sanitize_driver.c"), "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="3")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    report = r.stdout[-2000:] + r.stderr[-4000:]
    assert r.returncode == 0, report
    assert r.stdout.strip().endswith("OK"), report
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, report
