"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(make -C oracle selftest: a host-only gcc build of every oracle source plus
oracle/selftest.c) runs its entry points on small inputs -- Philox, cross
sections, both multigrid solvers, the three step loops, the Monte-Carlo loop
with pair re-injection, the steady-wall pairing, source / reduction /
weighted deposit, the wall surface physics and dump I/O -- with no sanitizer
report, no leak and every result finite."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                    reason="needs gcc and make")
def test_oracle_under_address_and_ub_sanitizers(tmp_path):
    b = subprocess.run(["make", "-s", "-C", ORACLE, "selftest"], capture_output=True, text=True)
    if b.returncode != 0 and "sanitize" in b.stderr and "cannot find" in b.stderr:
        pytest.skip("the sanitizer runtimes are not installed: " + b.stderr[-300:])
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(ORACLE, "_san", "selftest")], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "selftest: ok" in r.stdout
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]
