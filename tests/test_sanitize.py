"""AddressSanitizer + UndefinedBehaviorSanitizer runs on the CPU: the oracle
(oracle/oracle.c, every entry on a small jittered lattice) and the host side of
the C ABI (every .hip source compiled host-only, plus the lattice generator),
built by tests/sanitize/Makefile.  A sanitizer report, a leak or a failed
check ends the driver with a non-zero status."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _build(target):
    if shutil.which("gcc") is None or not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("toolchain absent")
    jobs = str(min(8, os.cpu_count() or 1))
    r = subprocess.run(["make", "-C", SAN, "-j" + jobs, os.path.join(SAN, "_build", target)],
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return os.path.join(SAN, "_build", target)


@pytest.mark.parametrize("target", ["oracle_driver", "host_abi"])
def test_sanitized_driver_is_clean(target):
    exe = _build(target)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-5000:]
    assert r.stdout.strip().endswith(": ok")
