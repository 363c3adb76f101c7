"""The reference's own user programs (exe/main.cpp and exe/benchmark.cpp of
Tom-Olsen/2dRadiation) compile and link unchanged against the drop-in C++
classes: `make -C tests/cpp compat` builds them where they lie in the
reference checkout.  CPU only (compile and link); skipped where the reference
checkout is absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = os.environ.get("RT2D_REFERENCE", "/root/reference")


@pytest.mark.skipif(not os.path.exists(os.path.join(REFERENCE, "exe", "main.cpp")),
                    reason="reference checkout not present")
def test_reference_user_programs_compile_unchanged():
    cpp = os.path.join(ROOT, "tests", "cpp")
    r = subprocess.run(["make", "-s", "-C", cpp, "compat", f"REFERENCE={REFERENCE}"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    for exe in ("main", "benchmark"):
        path = os.path.join(cpp, "_build", "compat", exe)
        assert os.path.exists(path) and os.access(path, os.X_OK), exe
        # the program is the reference's file, not a copy
        src = os.path.join(cpp, "_build", "compat", "exe", exe + ".cpp")
        assert os.path.realpath(src) == os.path.realpath(os.path.join(REFERENCE, "exe", exe + ".cpp"))
