"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

* The library's host-only parsers and writers (io.cpp: parameter cards, SU
  files; tuning.cpp: the JSON tuning cache) built with g++ -fsanitize=
  address,undefined and driven over the reference's cards, every truncation,
  written / merged SU files and byte-corrupted cache files
  (tests/sanitize/host_driver.cpp), leak checking on.
* The CPU oracle built the same way (gcc) and every entry point called on a
  small case in a subprocess with the ASan runtime preloaded
  (tests/sanitize/oracle_calls.py).
Any sanitizer report ends the process with a non-zero status.
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
ENV = dict(os.environ, UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="host compilers missing")


def test_host_io_and_tuning_cache(tmp_path):
    subprocess.run(["make", "-s", "-C", SAN, "host", f"OUT={tmp_path}"], check=True)
    work = tmp_path / "work"
    work.mkdir()
    r = subprocess.run([str(tmp_path / "host_driver"), os.path.join(HERE, "golden"), str(work)],
                       env=dict(ENV, ASAN_OPTIONS="detect_leaks=1"), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout and "runtime error" not in r.stderr


def test_oracle_entry_points(tmp_path):
    subprocess.run(["make", "-s", "-C", SAN, "oracle", f"OUT={tmp_path}"], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(ENV, SEISFD_ORACLE_LIB=str(tmp_path / "liboracle_asan.so"), LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0", OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, os.path.join(SAN, "oracle_calls.py")], env=env, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "oracle calls ok" in r.stdout and "runtime error" not in r.stderr
