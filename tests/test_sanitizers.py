"""AddressSanitizer / UndefinedBehaviorSanitizer runs on the CPU (oracle/Makefile, target
`sanitize`): the oracle driven through every scheme, boundary and option combination on small
blocks (oracle/sanitize_oracle.c), and the host side of the C ABI -- nsgpu_capi.hip with the
sanitizers on its host code, linked to the product's kernel objects -- driven through its
argument and configuration validation (oracle/sanitize_capi.cpp).  Any finding aborts the
program; leaks are reported by LeakSanitizer."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

ORACLE = os.path.join(ROOT, "oracle")
LIB = os.path.join(ROOT, "3dnssolver_amd", "lib")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _run(target):
    r = subprocess.run(["make", "-s", "-C", ORACLE, target], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([os.path.join(ORACLE, target)], capture_output=True, text=True,
                       timeout=600, env=ENV)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    return r.stdout


@pytest.mark.skipif(not shutil.which("gcc"), reason="no C compiler")
def test_oracle_under_asan_ubsan():
    assert "0 failures" in _run("_san/oracle_san")


@pytest.mark.skipif(not (shutil.which("hipcc") and os.path.exists(os.path.join(LIB, "kernels_exact.o"))),
                    reason="hipcc or the product's kernel objects absent")
def test_capi_host_code_under_asan_ubsan():
    assert "0 failures" in _run("_san/capi_san")
