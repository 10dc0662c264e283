"""The library's host logic without a GPU: a build of libfdtd_hip's host
code that records kernel launches instead of running them
(tests/host_record), checked for the launch decomposition of every kernel
and the step_n schedule over many shapes, dtypes, Poynting modes and Bloch
settings; once plainly and once with AddressSanitizer and UBSan on the
host code (the reference's host loop it replaces: fdtd.py:162-175)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "host_record"))

import build_record  # noqa: E402


@pytest.fixture(scope="module")
def record_libs():
    if not (os.path.exists(build_record.PLAIN) and os.path.exists(build_record.ASAN)):
        build_record.build()
    return build_record.PLAIN, build_record.ASAN


def test_schedule_and_decomposition(record_libs):
    from tests.host_record import checks
    n = checks.run(record_libs[0])
    assert n > 7000


def test_schedule_under_address_sanitizer(record_libs):
    env = dict(os.environ, LD_PRELOAD=build_record.asan_runtime(),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "host_record", "checks.py"), record_libs[1], ROOT],
                       capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "host schedule checks:" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
