"""The gppKer / ffKer host programs: argument handling as the reference's
programs (BGW-Kernels/gppKer.cpp:164-168, FFKer_cori/ffKernel.cpp:305-325)
on the CPU, and their printed results on the GPU against the reference
program's golden outputs."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

BIN = os.path.join(ROOT, "fortkernels-_amd", "bin")


def _run(exe, *args, timeout=120):
    return subprocess.run([os.path.join(BIN, exe), *map(str, args)], capture_output=True, text=True,
                          timeout=timeout)


def _complex_lines(text, name):
    out = {}
    for m in re.finditer(rf"^{name}\[(\d+)\] = \(([^,]+),([^)]+)\)$", text, re.M):
        out[int(m.group(1))] = complex(float(m.group(2)), float(m.group(3)))
    return [out[i] for i in sorted(out)]


@pytest.mark.parametrize("args", [[], ["8", "2", "1024"], ["8", "2", "1024", "20", "5"], ["1", "1", "1", "1",
                                                                                         "--bogus"]])
def test_gppker_usage_exits_zero(args):
    r = _run("gppKer", *args)
    assert r.returncode == 0
    assert "The correct form of input is" in r.stdout


def test_ffker_usage_and_ngpown_check():
    r = _run("ffKer", 8, 2, 64)
    assert r.returncode == 0 and "Incorrect Parameters" in r.stdout
    r = _run("ffKer", 8, 2, 32, 64, 4, 3)         # ngpown > ncouls
    assert r.returncode == 0 and "ngpown cannot be greater than ncouls" in r.stdout


@pytest.mark.gpu
def test_gppker_matches_reference_program():
    d = np.load(os.path.join(GOLDEN, "gpp_main.npz"))
    for (nb, nv, C, npg), ref in zip(d["shapes"], d["achtemp"]):
        print(f"gppKer {nb} {nv} {C} {npg}", flush=True)
        r = _run("gppKer", nb, nv, C, npg)
        assert r.returncode == 0, r.stderr
        ach = _complex_lines(r.stdout, "achtemp")
        assert len(ach) == 3
        for g, w in zip(ach, ref):
            assert abs(g - w) <= 1e-13 * abs(w), (g, w)
        assert "********** Chrono Time Taken **********= " in r.stdout


@pytest.mark.gpu
def test_gppker_random_matches_library(fk, fko):
    """--random builds the seeded inputs on the device; the printed values
    equal the oracle's on the same inputs."""
    nb, nv, C, npg = 16, 2, 512, 20
    r = _run("gppKer", nb, nv, C, npg, "--random", 3, "--acht-n1")
    assert r.returncode == 0, r.stderr
    p = fko.gpp_random(nb, nv, C, C // npg, seed=3)
    ref_a, ref_s, ref_n1 = fko.gpp(p, want_n1loc=True)
    for name, ref in (("achtemp", ref_a), ("asxtemp", ref_s), ("acht_n1_loc", ref_n1)):
        got = np.array(_complex_lines(r.stdout, name))
        assert got.shape == ref.shape
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref)), name


@pytest.mark.gpu
def test_ffker_matches_reference_program():
    d = np.load(os.path.join(GOLDEN, "ff_main.npz"))
    for (nb, nv, C, G, nf, nfe), ref in zip(d["shapes"], d["achsDtemp"]):
        r = _run("ffKer", nb, nv, C, G, nf, nfe)
        assert r.returncode == 0, r.stderr
        m = re.search(r"^achsDtemp \(full precision\) = \(([^,]+), ([^)]+)\)$", r.stdout, re.M)
        got = complex(float(m.group(1)), float(m.group(2)))
        assert abs(got - ref) <= 1e-13 * abs(ref)
        assert re.search(r"^achsDtemp = \( -?\d+\.\d{6}, -?\d+\.\d{6}\) $", r.stdout, re.M)
