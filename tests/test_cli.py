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


@pytest.mark.parametrize("args", [[], ["8", "2", "1024"], ["8", "2", "1024", "20", "1", "5"],
                                  ["1", "1", "1", "1", "--bogus"]])
def test_gppker_usage_exits_zero(args):
    r = _run("gppKer", *args)
    assert r.returncode == 0
    assert "The correct form of input is" in r.stdout


@pytest.mark.parametrize("args", [["8", "2", "1024", "20", "0"], ["8", "2", "1024", "20", "2"],
                                  ["8", "2", "1024", "20", "1", "--mixed"]])
def test_gppker_five_number_form_needs_gppsum_one(args):
    """The five-number form is gppKerSeq.cpp's command line; only gppsum = 1
    (FP64) is built, anything else stops before any device work."""
    r = _run("gppKer", *args)
    assert r.returncode == 2 and "gppsum = 1" in r.stderr


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
def test_gppker_gppsum_matches_reference_program():
    """gppKer <nb> <nv> <ncouls> <npg> 1 prints what the reference's
    gppKerSeq program prints for gppsum = 1 (its constant inputs)."""
    d = np.load(os.path.join(GOLDEN, "gppsum_main.npz"))
    for (nb, nv, C, npg), ref in zip(d["shapes"], d["achtemp"]):
        print(f"gppKer {nb} {nv} {C} {npg} 1", flush=True)
        r = _run("gppKer", nb, nv, C, npg, 1)
        assert r.returncode == 0, r.stderr
        ach = _complex_lines(r.stdout, "achtemp")
        assert len(ach) == 3
        for g, w in zip(ach, ref):
            assert abs(g - w) <= 1e-13 * abs(w), (g, w)


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
def test_gppker_achstemp(fko):
    """--achstemp prints the Fortran static-CH remainder; with the constant
    inputs it is -0.5 nb ngpown ncouls (0.25+0.25i)."""
    nb, nv, C, npg = 16, 2, 1024, 20
    r = _run("gppKer", nb, nv, C, npg, "--achstemp")
    assert r.returncode == 0, r.stderr
    m = re.search(r"^achstemp = \(([^,]+),([^)]+)\)$", r.stdout, re.M)
    got = complex(float(m.group(1)), float(m.group(2)))
    want = -0.5 * nb * (C // npg) * C * (0.25 + 0.25j)
    assert abs(got - want) <= 1e-13 * abs(want)
    r = _run("gppKer", nb, nv, C, npg, "--random", 2, "--achstemp")
    assert r.returncode == 0, r.stderr
    m = re.search(r"^achstemp = \(([^,]+),([^)]+)\)$", r.stdout, re.M)
    p = fko.gpp_random(nb, nv, C, C // npg, seed=2)
    want = fko.gpp_achstemp(p)
    assert abs(complex(float(m.group(1)), float(m.group(2))) - want) <= 1e-12 * abs(want)


@pytest.mark.gpu
def test_ffker_matches_reference_program(fko):
    """achsDtemp as the C++ program prints it (frequency index 1); asxDtemp,
    achDtemp_cor and the rest follow the Fortran program, checked against the
    oracle's canonical variant on the same constant inputs."""
    d = np.load(os.path.join(GOLDEN, "ff_main.npz"))
    for (nb, nv, C, G, nf, nfe), ref in zip(d["shapes"], d["achsDtemp"]):
        nb, nv, C, G, nf, nfe = (int(x) for x in (nb, nv, C, G, nf, nfe))
        print(f"ffKer {nb} {nv} {C} {G} {nf} {nfe}", flush=True)
        r = _run("ffKer", nb, nv, C, G, nf, nfe)
        assert r.returncode == 0, r.stderr
        m = re.search(r"^achsDtemp \(full precision\) = \(([^,]+), ([^)]+)\)$", r.stdout, re.M)
        got = complex(float(m.group(1)), float(m.group(2)))
        assert abs(got - ref) <= 1e-13 * abs(ref)
        for name in ("achsDtemp", "asxDtemp", "achDtemp_cor"):
            assert re.search(rf"^{name} = \( -?\d+\.\d{{6}}, -?\d+\.\d{{6}}\) $", r.stdout, re.M), name
        assert re.search(r"^ Answer: \(", r.stdout, re.M)
        ekq, grid, pref, brd = fko.ff_program_axes(nb, nf)
        z = np.full((nb, C), 0.5 + 0.5j)
        inv = np.arange(G, dtype=np.int32)
        ind = np.arange(C + 1, dtype=np.int32)
        ind[C] = C - 1
        p = fko.FfFullProblem(nb, nv, C, G, nf, nfe, z, z.copy(), np.full((nf, G, C), 0.5 + 0.5j),
                              np.full((nf, G, C), 0.5 - 0.5j), np.ones(C), inv, ind, ekq, grid, pref, brd)
        want = fko.ff_full(p)
        for name in fko.FF_OUTPUTS:
            got = np.array(_complex_lines(r.stdout, name))
            assert got.shape == want[name].shape, name
            assert np.max(np.abs(got - want[name])) <= 1e-12 * max(np.max(np.abs(want[name])), 1e-300), name


@pytest.mark.gpu
def test_ffker_cpp_semantics_prints_the_cpp_program(fko):
    """--semantics 2 prints what the C++ ffKernel program prints
    (ff_main.npz: its asxDtemp[0] and achDtemp_cor[0] at %f, achsDtemp at
    full precision), and every output at full precision as the oracle's
    variant 1 with schDt_matrix zero."""
    d = np.load(os.path.join(GOLDEN, "ff_main.npz"))
    for (nb, nv, C, G, nf, nfe), achs, asx, cor in zip(d["shapes"], d["achsDtemp"], d["asxDtemp"],
                                                        d["achDtemp_cor"]):
        nb, nv, C, G, nf, nfe = (int(x) for x in (nb, nv, C, G, nf, nfe))
        print(f"ffKer {nb} {nv} {C} {G} {nf} {nfe} --semantics 2", flush=True)
        r = _run("ffKer", nb, nv, C, G, nf, nfe, "--semantics", 2)
        assert r.returncode == 0, r.stderr
        printed = {}
        for name in ("achsDtemp", "asxDtemp", "achDtemp_cor"):
            m = re.search(rf"^{name} = \( (-?\d+\.\d{{6}}), (-?\d+\.\d{{6}})\) $", r.stdout, re.M)
            printed[name] = complex(float(m.group(1)), float(m.group(2)))
        for name, ref in (("achsDtemp", achs), ("asxDtemp", asx), ("achDtemp_cor", cor)):
            assert abs(printed[name] - ref) <= 1e-6 * max(1.0, abs(ref)), (name, printed[name], ref)
        ekq, grid, pref, brd = fko.ff_program_axes(nb, nf)
        z = np.full((nb, C), 0.5 + 0.5j)
        inv = np.arange(G, dtype=np.int32)
        ind = np.arange(C + 1, dtype=np.int32)
        ind[C] = C - 1
        p = fko.FfFullProblem(nb, nv, C, G, nf, nfe, z, z.copy(), np.full((nf, G, C), 0.5 + 0.5j),
                              np.full((nf, G, C), 0.5 - 0.5j), np.ones(C), inv, ind, ekq, grid, pref, brd,
                              achs_freq=1)
        want = fko.ff_full(p, variant=1)
        for name in fko.FF_OUTPUTS:
            got = np.array(_complex_lines(r.stdout, name))
            scale = max(np.max(np.abs(want[name])), 1e-300)
            assert np.max(np.abs(got - want[name])) <= 1e-12 * scale, name


@pytest.mark.gpu
def test_ffker_random_matches_oracle(fko):
    """--random fills aqs*, I_epsR, I_epsA, vcoul on the device with the seeded
    generator the oracle shares; every printed output equals the oracle's."""
    nb, nv, C, G, nf, nfe = 16, 4, 24, 8, 6, 10
    r = _run("ffKer", nb, nv, C, G, nf, nfe, "--random", 3, "--achs-freq", 0)
    assert r.returncode == 0, r.stderr
    want = fko.ff_full(fko.ff_full_random(nb, nv, C, G, nf, nfe, seed=3))
    for name in fko.FF_OUTPUTS:
        got = np.array(_complex_lines(r.stdout, name))
        assert np.max(np.abs(got - want[name])) <= 1e-12 * np.max(np.abs(want[name])), name
    m = re.search(r"^achsDtemp \(full precision\) = \(([^,]+), ([^)]+)\)$", r.stdout, re.M)
    assert abs(complex(float(m.group(1)), float(m.group(2))) - want["achsDtemp"]) <= 1e-12 * abs(want["achsDtemp"])
    r2 = _run("ffKer", nb, nv, C, G, nf, nfe, "--random", 3, "--achs-freq", 0, "--mixed")
    assert r2.returncode == 0, r2.stderr
    got = np.array(_complex_lines(r2.stdout, "achDtemp"))
    assert np.max(np.abs(got - want["achDtemp"])) <= 1e-5 * np.max(np.abs(want["achDtemp"]))


@pytest.mark.gpu
def test_ffker_random_port_semantics_and_bad_value(fko):
    """--semantics 1 on the seeded inputs: every printed output equals the
    oracle's variant 1 with the oracle's own contraction as schDt_matrix;
    an unknown semantics value is the library's argument error."""
    nb, nv, C, G, nf, nfe = 16, 4, 24, 8, 6, 10
    r = _run("ffKer", nb, nv, C, G, nf, nfe, "--random", 5, "--semantics", 1)
    assert r.returncode == 0, r.stderr
    p = fko.ff_full_random(nb, nv, C, G, nf, nfe, seed=5)
    p.achs_freq = 1
    want = fko.ff_full(p, variant=1, schDt_in=fko.ff_full(p)["schDt"])
    for name in fko.FF_OUTPUTS:
        got = np.array(_complex_lines(r.stdout, name))
        scale = max(np.max(np.abs(want[name])), 1e-300)
        assert np.max(np.abs(got - want[name])) <= 1e-12 * scale, name
    bad = _run("ffKer", nb, nv, C, G, nf, nfe, "--semantics", 7)
    assert bad.returncode != 0 and "semantics" in bad.stderr


@pytest.mark.gpu
def test_ffker_contraction_mode(fko):
    d = np.load(os.path.join(GOLDEN, "ff_main.npz"))
    (nb, nv, C, G, nf, nfe), ref = d["shapes"][0], d["achsDtemp"][0]
    r = _run("ffKer", nb, nv, C, G, nf, nfe, "--contraction", "--both")
    assert r.returncode == 0, r.stderr
    m = re.search(r"^achsDtemp \(full precision\) = \(([^,]+), ([^)]+)\)$", r.stdout, re.M)
    assert abs(complex(float(m.group(1)), float(m.group(2))) - ref) <= 1e-13 * abs(ref)
    assert "schDt[I_epsA]" in r.stdout
