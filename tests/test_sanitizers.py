"""The oracle and libhhg_amd's host code under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only): oracle/Makefile's `sanitize` driver
runs every model, pulse, relaxation, method and controller failure path;
tests/sanitize/host_check.cpp runs the host utilities and job validation.
Any sanitizer report or failed expectation fails the test."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build_and_run(make_dir, target, exe):
    r = subprocess.run(["make", "-s", target], cwd=make_dir, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    report = run.stdout + run.stderr
    assert run.returncode == 0, report
    assert "Sanitizer" not in report and "runtime error" not in report, report
    assert "0 failed expectations" in report


def test_oracle_under_asan_and_ubsan():
    d = os.path.join(ROOT, "oracle")
    _build_and_run(d, "sanitize", os.path.join(d, "_san", "oracle_check"))


def test_host_code_under_asan_and_ubsan():
    d = os.path.join(ROOT, "tests", "sanitize")
    _build_and_run(d, "all", os.path.join(d, "_build", "host_check"))
