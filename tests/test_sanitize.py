"""Host-code sanitizer run: the CPU oracle and the library's host graph code (readers, binary
CSR cache, host generators) built with AddressSanitizer + UndefinedBehaviorSanitizer by
tests/san/Makefile and run over the fixtures and hostile inputs; any report fails the test."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="no host compiler")
def test_host_code_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-C", os.path.join(HERE, "san"), "run"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "oracle_san: " in r.stdout and "graphio_san: ok" in r.stdout, r.stdout[-2000:]
    assert "runtime error" not in r.stdout and "AddressSanitizer" not in r.stdout, r.stdout[-4000:]
