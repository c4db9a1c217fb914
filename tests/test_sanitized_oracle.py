"""The CPU oracle and the host emulation of the lower-star kernel, built with
AddressSanitizer and UndefinedBehaviorSanitizer (g++, host code only) and run
over small volumes with boundaries, ties and budget-bound streamlines
(tests/sanitize/driver.cpp): the checkers themselves must be free of
out-of-bounds accesses and undefined arithmetic."""
import os
import subprocess

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


def test_oracle_and_emulation_are_clean_under_sanitizers():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(HERE, "sanitized")], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "ok" in r.stdout
    assert "runtime error" not in r.stderr, r.stderr[-4000:]
