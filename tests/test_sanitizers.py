"""CPU builds with AddressSanitizer and UndefinedBehaviorSanitizer
(tests/sanitize/Makefile): the library's host-side format code (BED parsing
over host threads, meta.json, jump tables, result JSON) driven by
tests/sanitize/host_driver.cpp, and the oracle under its own CPU tests; and
with ThreadSanitizer: the library's host worker pool driven by
tests/sanitize/pool_stress.cpp."""
import gzip
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
OUT = os.path.join(SAN, "_build")


def gcc_lib(name):
    return subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-s", "-C", SAN], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return OUT


def san_env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    return env


def test_host_format_code_under_sanitizers(built, tmp_path):
    rng = np.random.default_rng(5)
    (tmp_path / "a.bed").write_text("# c\n\nchr1\t10\t20\n30 40\n50\t60\tchr2\tname\t7\nchrX 5 9 x 1 2\n")
    with gzip.open(tmp_path / "b.bed.gz", "wt") as f:
        f.write("chr1\t100\t200\nchr3\t1\t2")  # no final newline
    s = rng.integers(0, 10**8, 900_000)
    big = "\n".join(f"chr{1 + i % 22}\t{a}\t{a + 1 + i % 997}" for i, a in enumerate(s.tolist()))
    (tmp_path / "big.bed").write_text(big + "\n")
    (tmp_path / "bad.bed").write_text("chr1 1 2\nchr1 9 3\n")
    (tmp_path / "empty.bed").write_text("")
    files = [str(tmp_path / n) for n in ("a.bed", "b.bed.gz", "big.bed", "bad.bed", "empty.bed")]
    r = subprocess.run([os.path.join(built, "host_driver"), str(tmp_path / "work"), *files],
                       capture_output=True, text=True, env=san_env(), timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout.splitlines()
    assert out[0] == "a.bed: 4 intervals, 4 shards"
    assert out[1] == "b.bed.gz: 2 intervals, 2 shards"
    assert out[2] == "big.bed: 900000 intervals, 22 shards"
    assert out[3] == "bad.bed: error: Line 2: start (9) must be less than end (3)"
    assert out[4] == "empty.bed: 0 intervals, 0 shards"
    assert out[-1] == "ok" and "runtime error" not in r.stderr


def test_oracle_under_sanitizers(built):
    env = san_env()
    env["LD_PRELOAD"] = ":".join([gcc_lib("libasan.so"), gcc_lib("libubsan.so")])
    env["CHK_ORACLE_LIB"] = os.path.join(built, "liboracle.so")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                        "tests/test_oracle.py", "tests/test_golden.py"],
                       cwd=ROOT, capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


def test_host_pool_under_thread_sanitizer(built):
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=1"
    r = subprocess.run([os.path.join(built, "pool_stress")], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "bad=0" and "ThreadSanitizer" not in r.stderr
