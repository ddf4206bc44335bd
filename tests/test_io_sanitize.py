"""The instance-file reader and writer (fortkernels-_amd/csrc/fk_io.cpp, host
code of the product library) under AddressSanitizer and
UndefinedBehaviorSanitizer: tests/sanitize_io_driver.cpp writes GPP and FF
instances, reads them back, opens about 12000 truncated and header-damaged
copies (every accepted one is checked against the documented layout by the
driver's own parser and read through), and reads instance lists with long
lines, CRLF ends and comments.  CPU only; skips when g++ has no sanitizer
runtime."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_instance_files_clean_under_asan_ubsan(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "sanitize_io"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           "-o", str(exe), os.path.join(ROOT, "fortkernels-_amd", "csrc", "fk_io.cpp"),
           os.path.join(HERE, "sanitize_io_driver.cpp")]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if b.returncode != 0 and ("asan" in b.stderr or "ubsan" in b.stderr):
        pytest.skip(f"g++ cannot build with sanitizers here: {b.stderr.strip()[:200]}")
    assert b.returncode == 0, b.stderr
    work = tmp_path / "scratch"
    work.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), str(work)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, f"stdout:\n{r.stdout}\nstderr:\n{r.stderr}"
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    assert "ERROR: LeakSanitizer" not in r.stderr, r.stderr
    out = r.stdout
    assert "sanitize_io_driver: all cases passed" in out, out
    for kind in ("GPP complex128", "GPP complex64", "FF"):
        assert f"{kind}: " in out and "accepted (each checked and read)" in out, out
    assert "instance lists: ok" in out
