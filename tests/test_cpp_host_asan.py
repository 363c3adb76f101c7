"""Memory checks of the host C++ classes on the CPU.

The reference-style driver (tests/cpp/radiation_driver.cpp) and the drop-in
classes (2dradiation_amd/cpp) are built with AddressSanitizer and UBSan over
a host-only stand-in of the C ABI (tests/cpp/rt_host_stub.cpp) that reads
and writes every element of the caller's buffers at the sizes include/rt2d.h
documents.  Every driver mode, construction to teardown, must run clean:
a heap overflow, a too-small buffer handed to the C ABI, a use after free or
undefined behaviour in the host code ends the process with a report.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
EXE = os.path.join(CPP, "_build", "radiation_driver_asan")
HELPERS = os.path.join(CPP, "_build", "radiation_helpers_asan")


@pytest.fixture(scope="module")
def driver():
    subprocess.run(["make", "-s", "-C", CPP, "asan"], check=True)
    return EXE


@pytest.mark.parametrize("mode", ["flat", "curved", "frames", "vonmises", "adaptive", "adaptive_curved"])
def test_driver_host_code_is_memory_clean(driver, mode, tmp_path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=86",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([driver, mode, str(tmp_path / "out.bin")], cwd=tmp_path, env=env, capture_output=True,
                       text=True, timeout=600)
    report = r.stdout[-4000:] + r.stderr[-4000:]
    assert r.returncode == 0, report
    assert "AddressSanitizer" not in report and "runtime error" not in report, report


def test_radiation_helpers_by_identities(driver, tmp_path):
    """GetTempCoordinate, GetTemp2VelocityIF, GetFrequencyShift, IntensityAt,
    AverageF, InitialDataLFtoIF and NormalizationOverwrite on host buffers
    (tests/cpp/radiation_helpers_check.cpp), under the same sanitizers."""
    r = subprocess.run([HELPERS], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "radiation_helpers OK" in r.stdout, r.stdout[-4000:] + r.stderr[-4000:]
