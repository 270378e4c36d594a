"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
product's parameter tables and modular arithmetic (params.cpp,
modarith.hpp) and the oracle, built with g++ -fsanitize=address,undefined
and run over small and large shapes (tests/native/sanitize_main.cpp).  Any
sanitizer report, or a failed check, fails the test."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "native", "_build", "sanitize_main")


@pytest.fixture(scope="module")
def binary():
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I" + os.path.join(ROOT, "aoa_amd", "csrc"),
           os.path.join(HERE, "native", "sanitize_main.cpp"), os.path.join(HERE, "native", "modarith_check.cpp"),
           os.path.join(ROOT, "aoa_amd", "csrc", "params.cpp"), os.path.join(ROOT, "oracle", "oracle.cpp"),
           "-o", OUT + ".tmp"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


def test_host_code_is_clean_under_asan_and_ubsan(binary):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([binary], capture_output=True, text=True, timeout=900, env=env)
    print(r.stdout[-3000:])
    print(r.stderr[-6000:])
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert r.returncode == 0
    assert "0 failure(s)" in r.stdout
