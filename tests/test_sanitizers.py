"""The repository's CPU code under AddressSanitizer and
UndefinedBehaviorSanitizer: tests/sanitize/Makefile builds the oracle
(oracle/ifos_oracle.c) and the library's SU file code
(ifos3d-dev_amd/csrc/suio_host.cpp) with -fsanitize=address,undefined and
san_driver.cpp runs gradient shots (elastic, free surface, three relaxation
mechanisms, the time-domain and the consistent-adjoint gradients) and SU
file round trips through them; any report ends the run with a non-zero
status."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="needs gcc and g++")
def test_cpu_code_under_asan_and_ubsan(tmp_path):
    b = subprocess.run(["make", "-s", "-C", SAN], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(SAN, "lib", "san_driver"), str(tmp_path)], capture_output=True, text=True,
                       timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    assert "sanitized run: OK" in out
