"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle sanitize` builds the oracle's C sources together with
oracle/fko_sanitize.c (a driver that calls every entry point on small,
ragged problems, each array its own exact-size heap block) with
-fsanitize=address,undefined -fno-sanitize-recover=all.  The run must finish
with no finding, and the fixed-order extended-precision sums must give the
same bits with 1 and 3 OpenMP threads.

`make -C oracle tsan` builds the same driver with ThreadSanitizer (ROCm's
clang and LLVM's OpenMP runtime, whose Archer tool reports OpenMP's
synchronisation to TSan); the run must report no race.

Control programs with a planted error (a one-past-the-end heap read under
ASan, a signed overflow under UBSan, an unsynchronised `+=` in an OpenMP
loop under TSan) must each be stopped: that shows each runtime is active
here, so a clean oracle run means something.  Tests skip when their
toolchain is missing.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")
BINARY = os.path.join(ORACLE, "_san", "fko_sanitize")
ENV = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
CLANG = "/opt/rocm/llvm/bin/clang"
ARCHER = "/opt/rocm/llvm/lib/libarcher.so"
TSAN_ENV = dict(os.environ, OMP_TOOL_LIBRARIES=ARCHER, OMP_NUM_THREADS="4",
                TSAN_OPTIONS="ignore_noninstrumented_modules=1:halt_on_error=1:exitcode=66")


def _need_toolchain():
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc/make not available")


CONTROLS = {
    # a one-past-the-end heap read (argc = 1), ASan alone
    "address": ("#include <stdlib.h>\n"
                "int main(int argc, char **argv) {\n"
                "    (void)argv;\n"
                "    double *a = malloc(4 * sizeof(double));\n"
                "    for (int i = 0; i < 4; ++i) a[i] = i;\n"
                "    volatile double x = a[3 + argc];\n"
                "    free(a);\n"
                "    return (int)x;\n"
                "}\n", "heap-buffer-overflow"),
    # a signed overflow, UBSan alone
    "undefined": ("#include <limits.h>\n"
                  "int main(int argc, char **argv) {\n"
                  "    (void)argv;\n"
                  "    volatile int x = INT_MAX;\n"
                  "    int y = x + argc;\n"
                  "    return y == 0;\n"
                  "}\n", "signed integer overflow"),
}


@pytest.mark.parametrize("kind", sorted(CONTROLS))
def test_sanitizer_runtime_catches_a_planted_error(tmp_path, kind):
    _need_toolchain()
    code, message = CONTROLS[kind]
    src = tmp_path / "control.c"
    src.write_text(code)
    exe = tmp_path / "control"
    flags = ["-O1", "-g", f"-fsanitize={kind}", "-fno-sanitize-recover=all"]
    b = subprocess.run(["gcc", *flags, "-o", str(exe), str(src)], capture_output=True, text=True)
    if b.returncode != 0:
        pytest.skip(f"gcc cannot build with -fsanitize={kind} here: {b.stderr.strip()[:200]}")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=ENV, timeout=60)
    assert r.returncode != 0
    assert message in r.stderr, r.stderr


def test_oracle_clean_under_asan_ubsan():
    _need_toolchain()
    b = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True, timeout=600)
    if b.returncode != 0 and ("asan" in b.stderr or "ubsan" in b.stderr or "sanitize" in b.stderr):
        pytest.skip(f"gcc cannot build with sanitizers here: {b.stderr.strip()[:200]}")
    assert b.returncode == 0, b.stderr
    r = subprocess.run([BINARY], capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0, f"stdout:\n{r.stdout}\nstderr:\n{r.stderr}"
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr
    assert "ERROR: LeakSanitizer" not in r.stderr, r.stderr
    assert r.stdout.strip().endswith("fko_sanitize: all cases passed"), r.stdout
    # every entry point was driven
    for name in ("fko_gpp ", "fko_ff ", "fko_ff_full ", "fko_gpp_point", "rejected index maps"):
        assert name in r.stdout, name


def _need_tsan():
    if not (os.path.exists(CLANG) and os.path.exists(ARCHER)) or shutil.which("make") is None:
        pytest.skip("clang with LLVM's OpenMP runtime and Archer not available")


def test_tsan_catches_a_planted_race(tmp_path):
    _need_tsan()
    src = tmp_path / "race.c"
    src.write_text("#include <stdio.h>\n"
                   "int main(void) {\n"
                   "    long sum = 0;\n"
                   "#pragma omp parallel for num_threads(4)\n"
                   "    for (int i = 0; i < 100000; ++i) sum += i;   /* no reduction clause */\n"
                   "    printf(\"%ld\\n\", sum);\n"
                   "    return 0;\n"
                   "}\n")
    exe = tmp_path / "race"
    llvmlib = os.path.dirname(ARCHER)
    b = subprocess.run([CLANG, "-O1", "-g", "-fopenmp", "-fsanitize=thread", "-o", str(exe), str(src),
                        f"-Wl,-rpath,{llvmlib}"], capture_output=True, text=True)
    if b.returncode != 0:
        pytest.skip(f"clang cannot build with -fsanitize=thread here: {b.stderr.strip()[:200]}")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=TSAN_ENV, timeout=120)
    assert r.returncode != 0
    assert "ThreadSanitizer: data race" in r.stderr, r.stderr


def test_oracle_clean_under_tsan():
    _need_tsan()
    b = subprocess.run(["make", "-s", "-C", ORACLE, "tsan"], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr
    r = subprocess.run([os.path.join(ORACLE, "_san", "fko_tsan")], capture_output=True, text=True,
                       env=TSAN_ENV, timeout=600)
    assert "Archer" not in r.stderr or "please export" not in r.stderr, r.stderr
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr
    assert r.returncode == 0, f"stdout:\n{r.stdout}\nstderr:\n{r.stderr}"
    assert r.stdout.strip().endswith("fko_sanitize: all cases passed"), r.stdout
