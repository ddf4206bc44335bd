"""Write the golden vectors under tests/golden/ from the compiled reference.

Run where the reference checkout exists (its sources are compiled by
`make -C oracle ref` into oracle/_ref/); the GPU tests only read the
fixtures this script writes.

Fixtures (numpy .npz, no pickles; inputs and expected outputs only):
  gpp_<case>.npz   inputs + achtemp/asxtemp/acht_n1_loc + per-column
                   scht/ssxt sums written by oracle/_ref/ref_gpp_harness,
                   i.e. by the reference's flagOCC_solver/ssxt_scht_solver
                   (BGW-Kernels/gppKer.cpp:19-65,130-159)
  gpp_points.npz   per-point ssx/sch of the reference's ssxt_scht_solver
  gpp_main.npz     achtemp printed by the reference's own gppKer program
                   (BGW-Kernels/gppKer.cpp main, constant inputs) at small
                   argv shapes, OMP_NUM_THREADS=1
  ff_achs.npz      achsDtemp of the reference's achsDtemp_Kernel
                   (FFKer_cori/ffKernel.cpp:107-131, frequency index 1)
  ff_main.npz      achsDtemp, asxDtemp[0] and achDtemp_cor[0] printed by
                   the reference's ffKernel program (constant inputs)
  gpp_achstemp.npz GPP inputs + the static-CH remainder achstemp of the
                   Fortran GPP program (gppkernel.deslippe.09.30.2014.f90:
                   192-230), written by the reference's achsDtemp_Kernel
                   (FFKer_cori/ffKernel.cpp:107-131), which computes the same
                   sum, run with the GPP I_eps at its frequency index 1
  ff_full_<case>.npz  inputs + achsDtemp, asxDtemp, achDtemp, ach2Dtemp,
                   achDtemp_cor, achDtemp_corb written by
                   oracle/_ref/ref_ff_harness full, i.e. by the reference's
                   achsDtemp_Kernel, asxDtemp_Kernel, achDtemp_Kernel and
                   achDtemp_cor_Kernel (FFKer_cori/ffKernel.cpp:10-300) called
                   in the order of its main, one thread

  gppsum_<case>.npz  inputs + achtemp/asxtemp/acht_n1_loc of the gppsum = 1
                   band loop, written by oracle/_ref/ref_gppseq_harness, i.e.
                   by the reference's flagOCC_solver / noflagOCC_solver of
                   gppKerSeq.cpp (:76-261) called as its main calls them
  gppsum_main.npz  achtemp printed by the reference's gppKerSeq program
                   (gppKerSeq.cpp main, constant inputs, gppsum = 1)

Usage: python tests/golden/make_golden.py [gppsum | ff_full case ...]
(with `gppsum`, only the gppsum fixtures; with ff_full case names, only
those ff_full_<case>.npz fixtures are written)
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import fko  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref")


def run_gpp_harness(p: fko.GppProblem):
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([p.nb, p.nv, p.ncouls, p.ngpown, p.nw], np.int32).tofile(f)
            np.array([p.occ], np.float64).tofile(f)
            np.ascontiguousarray(p.wx, np.float64).tofile(f)
            for a in (p.aqsm, p.aqsn, p.I_eps, p.wtilde):
                np.ascontiguousarray(a, np.complex128).tofile(f)
            np.ascontiguousarray(p.vcoul, np.float64).tofile(f)
            np.ascontiguousarray(p.inv_igp_index, np.int32).tofile(f)
            np.ascontiguousarray(p.indinv, np.int32).tofile(f)
        subprocess.run([os.path.join(REF_BIN, "ref_gpp_harness"), fin, fout], check=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
        out = np.fromfile(fout, np.complex128)
    nw, nb, G = p.nw, p.nb, p.ngpown
    o = 0
    ach = out[o:o + nw]; o += nw
    asx = out[o:o + nw]; o += nw
    n1loc = out[o:o + nb]; o += nb
    scht = out[o:o + nb * G * nw].reshape(nb, G, nw); o += nb * G * nw
    ssxt = out[o:o + nb * G * nw].reshape(nb, G, nw); o += nb * G * nw
    assert o == out.size
    return ach, asx, n1loc, scht, ssxt


def save_gpp(name: str, p: fko.GppProblem, **extra):
    ach, asx, n1loc, scht, ssxt = run_gpp_harness(p)
    np.savez_compressed(
        os.path.join(HERE, f"gpp_{name}.npz"),
        nb=p.nb, nv=p.nv, ncouls=p.ncouls, ngpown=p.ngpown, occ=p.occ,
        wx=p.wx, aqsm=p.aqsm, aqsn=p.aqsn, I_eps=p.I_eps, wtilde=p.wtilde, vcoul=p.vcoul,
        inv_igp_index=p.inv_igp_index, indinv=p.indinv,
        achtemp=ach, asxtemp=asx, acht_n1_loc=n1loc, scht=scht, ssxt=ssxt, **extra)
    print(f"gpp_{name}: achtemp={ach} asxtemp={asx}")


def case_random():
    return fko.gpp_random(6, 2, 96, 7, nw=3, seed=11)


def case_nearpole():
    """wtilde.re ~ U(2.5,5.5) puts points within 0.5 of wx = 3,4,5 (branch 2),
    plus points within |wt|/500 of a pole (|delw|^2 >= limitone)."""
    p = fko.gpp_random(5, 3, 80, 6, nw=3, seed=12, wtilde_re=(2.5, 5.5))
    w = p.wtilde.copy()
    w[0, 3] = 3.0 + 1e-4 + 1e-5j
    w[1, 7] = 4.0 - 2e-3 + 0.0j
    w[2, 11] = 5.0 + 5e-3 - 1e-3j
    w[3, 0] = 3.2 + 0.4j
    p.wtilde = w
    return p


def case_branch3():
    """Small wx so that |wdiff|^2 <= limittwo with |delw|^2 <= to1 (branch 3)."""
    p = fko.gpp_random(4, 4, 64, 5, nw=3, seed=13, wtilde_re=(-0.05, 0.3))
    p.wx = np.array([0.2, 0.35, 3.0])
    w = p.wtilde.copy()
    rng = np.random.default_rng(13)
    idx = rng.choice(w.size, 60, replace=False)
    w.flat[idx] = (rng.uniform(-2e-4, 2e-4, 60) + 1j * rng.uniform(-2e-4, 2e-4, 60))
    p.wtilde = w
    return p


def case_skipmap():
    """C++ map (i+1)*C/G with the indigp == C clamp, plus skipped columns
    (negative indinv entries, BGW-Kernels/gppKer.cpp:322-328)."""
    p = fko.gpp_random(4, 1, 64, 9, nw=4, seed=14, map_kind="cpp")
    ind = p.indinv.copy()
    ind[p.inv_igp_index[2]] = -1
    ind[p.inv_igp_index[5]] = -7
    p.indinv = ind
    return p


def points():
    """Per-point golden values: nb=1, C=1, aqsm=aqsn=1 so each column sum is the
    point's ssx/sch exactly (flagOCC_solver with igmax=1)."""
    rng = np.random.default_rng(21)
    K = 40
    wt = rng.uniform(-1, 6, K) + 1j * rng.uniform(-0.3, 0.3, K)
    wt[:8] = np.array([3 + 1e-4j, 4.0001, 5 - 3e-3 + 1e-3j, 1e-4 + 1e-4j, 2.99, 3.3 + 0.1j, 0.19 + 0j, 4.2 - 0.2j])
    ie = rng.uniform(-0.5, 0.5, K) + 1j * rng.uniform(-0.5, 0.5, K)
    wx = np.array([0.2, 3.0, 4.0, 5.0, 1e-6])
    p = fko.GppProblem(1, 1, 1, K, wx, np.ones((1, 1), complex), np.ones((1, 1), complex),
                       ie.reshape(K, 1), wt.reshape(K, 1), np.ones(1), np.zeros(K, np.int32),
                       np.array([0, 0], np.int32))
    _, _, _, scht, ssxt = run_gpp_harness(p)
    np.savez_compressed(os.path.join(HERE, "gpp_points.npz"), wx=wx, wtilde=wt, I_eps=ie,
                        ssx=ssxt[0], sch=scht[0])
    print("gpp_points: written")


def reference_main():
    shapes = [(8, 2, 1024, 20), (16, 4, 512, 8), (6, 6, 256, 5)]
    rows = []
    for nb, nv, C, npg in shapes:
        out = subprocess.run([os.path.join(REF_BIN, "gppKer"), str(nb), str(nv), str(C), str(npg)],
                             check=True, capture_output=True, text=True,
                             env=dict(os.environ, OMP_NUM_THREADS="1")).stdout
        vals = re.findall(r"achtemp\[(\d)\] = \(([-\d.e+]+),([-\d.e+]+)\)", out)
        assert len(vals) == 3, out
        rows.append([complex(float(a), float(b)) for _, a, b in vals])
    np.savez_compressed(os.path.join(HERE, "gpp_main.npz"), shapes=np.array(shapes, np.int64),
                        achtemp=np.array(rows))
    print("gpp_main:", rows)


def run_ff_harness(p: fko.FfProblem):
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([p.nb, p.ncouls, p.ngpown, p.nfreq], np.int32).tofile(f)
            for a in (p.aqsm, p.aqsn, p.I_eps):
                np.ascontiguousarray(a, np.complex128).tofile(f)
            np.ascontiguousarray(p.vcoul, np.float64).tofile(f)
            np.ascontiguousarray(p.inv_igp_index, np.int32).tofile(f)
            np.ascontiguousarray(p.indinv[: p.ncouls], np.int32).tofile(f)
        subprocess.run([os.path.join(REF_BIN, "ref_ff_harness"), fin, fout], check=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
        return complex(*np.fromfile(fout, np.float64))


def ff_cases():
    p = fko.ff_random(5, 48, 6, 3, seed=31)
    achs = run_ff_harness(p)
    np.savez_compressed(os.path.join(HERE, "ff_achs.npz"), nb=p.nb, ncouls=p.ncouls, ngpown=p.ngpown,
                        nfreq=p.nfreq, aqsm=p.aqsm, aqsn=p.aqsn, I_eps=p.I_eps, vcoul=p.vcoul,
                        inv_igp_index=p.inv_igp_index, indinv=p.indinv, achsDtemp=achs)
    print("ff_achs:", achs)
    shapes = [(8, 2, 64, 32, 4, 3), (4, 2, 32, 16, 3, 2), (16, 4, 24, 8, 6, 10)]
    rows = {"achsDtemp": [], "asxDtemp": [], "achDtemp_cor": []}
    for s in shapes:
        out = subprocess.run([os.path.join(REF_BIN, "ffKernel")] + [str(v) for v in s], check=True,
                             capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="1")).stdout
        for k in rows:
            m = re.search(k + r" = \( ([-\d.]+), ([-\d.]+)\)", out)
            assert m, out
            rows[k].append(complex(float(m.group(1)), float(m.group(2))))
    np.savez_compressed(os.path.join(HERE, "ff_main.npz"), shapes=np.array(shapes, np.int64),
                        **{k: np.array(v) for k, v in rows.items()})
    print("ff_main:", rows)


def run_ff_full_harness(p: fko.FfFullProblem, schDt):
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([p.nb, p.nv, p.ncouls, p.ngpown, p.nfreq, p.nfreqeval], np.int32).tofile(f)
            np.array([p.freqevalmin, p.freqevalstep, p.occ, p.pref_zb], np.float64).tofile(f)
            for a in (p.aqsm, p.aqsn, p.I_epsR, p.I_epsA):
                np.ascontiguousarray(a, np.complex128).tofile(f)
            for a in (p.vcoul, p.ekq, p.dFreqGrid, p.pref):
                np.ascontiguousarray(a, np.float64).tofile(f)
            for a in (p.dFreqBrd, schDt):
                np.ascontiguousarray(a, np.complex128).tofile(f)
            np.ascontiguousarray(p.inv_igp_index, np.int32).tofile(f)
            np.ascontiguousarray(p.indinv[: p.ncouls], np.int32).tofile(f)
        subprocess.run([os.path.join(REF_BIN, "ref_ff_harness"), "full", fin, fout], check=True)
        v = np.fromfile(fout, np.complex128)
    ne = p.nfreqeval
    out = {"achsDtemp": v[0]}
    for i, k in enumerate(("asxDtemp", "achDtemp", "ach2Dtemp", "achDtemp_cor", "achDtemp_corb")):
        out[k] = v[1 + i * ne: 1 + (i + 1) * ne]
    return out


def ff_full_cases(only=None):
    """Seeded inputs for the C++ port's FF kernel functions: random band
    energies (both signs of wx for occupied bands), a non-uniform grid, an
    evaluation axis whose freqevalmin / freqevalstep the port truncates to
    int, and a seeded schDt_matrix (the port's main never fills it)."""
    rng = np.random.default_rng(2024)
    cases = {}
    p = fko.ff_full_random(14, 5, 40, 8, 6, 7, seed=41)
    p.ekq = rng.uniform(-12.0, 12.0, p.nb)
    p.dFreqGrid = np.cumsum(rng.uniform(0.5, 3.0, p.nfreq)) - 0.6
    p.dFreqGrid[0] = 0.0
    p.pref = rng.uniform(0.1, 0.6, p.nfreq)
    p.dFreqBrd = rng.uniform(-0.05, 0.05, p.nfreq) + 1j * rng.uniform(0.05, 0.2, p.nfreq)
    p.freqevalmin, p.freqevalstep, p.occ = 0.25, 1.5, 0.75
    cases["seeded"] = (p, (rng.uniform(-1, 1, (p.nb, p.nfreq)) + 1j * rng.uniform(-1, 1, (p.nb, p.nfreq))) * 50)
    q = fko.ff_full_random(12, 3, 24, 6, 5, 9, seed=42)     # the program's axes, schDt from the contraction
    cases["program_axes"] = (q, fko.ff_full(q)["schDt"])
    # a negative freqevalmin, which the port truncates toward zero (int(-1.75)
    # = -1), a non-uniform grid, schDt from the contraction
    rng2 = np.random.default_rng(2025)
    r = fko.ff_full_random(13, 4, 36, 7, 6, 8, seed=43)
    r.ekq = rng2.uniform(-12.0, 12.0, r.nb)
    r.dFreqGrid = np.cumsum(rng2.uniform(0.5, 3.0, r.nfreq)) - 0.4
    r.dFreqGrid[0] = 0.0
    r.pref = rng2.uniform(0.1, 0.6, r.nfreq)
    r.dFreqBrd = rng2.uniform(-0.05, 0.05, r.nfreq) + 1j * rng2.uniform(0.05, 0.2, r.nfreq)
    r.freqevalmin, r.freqevalstep, r.occ = -1.75, 2.5, 0.5
    cases["negative_axes"] = (r, fko.ff_full(r)["schDt"])
    for name, (p, schDt) in cases.items():
        if only and name not in only:
            continue
        out = run_ff_full_harness(p, schDt)
        np.savez_compressed(os.path.join(HERE, f"ff_full_{name}.npz"),
                            sizes=np.array([p.nb, p.nv, p.ncouls, p.ngpown, p.nfreq, p.nfreqeval], np.int64),
                            scalars=np.array([p.pref_zb, p.freqevalmin, p.freqevalstep, p.occ]),
                            aqsm=p.aqsm, aqsn=p.aqsn, I_epsR=p.I_epsR, I_epsA=p.I_epsA, vcoul=p.vcoul,
                            inv_igp_index=p.inv_igp_index, indinv=p.indinv, ekq=p.ekq, dFreqGrid=p.dFreqGrid,
                            pref=p.pref, dFreqBrd=p.dFreqBrd, schDt=schDt, **out)
        print(f"ff_full_{name}:", {k: np.round(v, 4) for k, v in out.items()})


def gpp_achstemp_case():
    """achsDtemp_Kernel sums -0.5 vcoul[igp] aqsn[n1][ig] conj(aqsm[n1][igp])
    I_epsR[1][my_igp][ig] over n1, my_igp, ig: the Fortran GPP achstemp when
    I_epsR[1] is the GPP I_eps.  Its column map has no skip rule, so the map
    here is a permutation."""
    p = fko.gpp_random(7, 2, 48, 6, seed=17)
    inv = np.array([5, 0, 3, 1, 4, 2], np.int32) * 8 + np.array([1, 7, 2, 0, 5, 3], np.int32)
    p.inv_igp_index = inv
    other = fko.fill_complex(p.ngpown * p.ncouls, 99, 3, (-0.5, 0.5), (-0.5, 0.5)).reshape(p.ngpown, p.ncouls)
    q = fko.FfProblem(p.nb, p.ncouls, p.ngpown, 2, p.aqsm, p.aqsn, np.stack([other, p.I_eps]), p.vcoul,
                      p.inv_igp_index, p.indinv)
    achs = run_ff_harness(q)
    np.savez_compressed(os.path.join(HERE, "gpp_achstemp.npz"), nb=p.nb, nv=p.nv, ncouls=p.ncouls,
                        ngpown=p.ngpown, wx=p.wx, aqsm=p.aqsm, aqsn=p.aqsn, I_eps=p.I_eps, wtilde=p.wtilde,
                        vcoul=p.vcoul, inv_igp_index=p.inv_igp_index, indinv=p.indinv, achstemp=achs)
    print("gpp_achstemp:", achs)


def run_gppseq_harness(p: fko.GppProblem, gppsum=1):
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([p.nb, p.nv, p.ncouls, p.ngpown, p.nw, gppsum], np.int32).tofile(f)
            np.array([p.occ], np.float64).tofile(f)
            np.ascontiguousarray(p.wx, np.float64).tofile(f)
            for a in (p.aqsm, p.aqsn, p.I_eps, p.wtilde):
                np.ascontiguousarray(a, np.complex128).tofile(f)
            np.ascontiguousarray(p.vcoul, np.float64).tofile(f)
            np.ascontiguousarray(p.inv_igp_index, np.int32).tofile(f)
            np.ascontiguousarray(p.indinv, np.int32).tofile(f)
        subprocess.run([os.path.join(REF_BIN, "ref_gppseq_harness"), fin, fout], check=True)
        out = np.fromfile(fout, np.complex128)
    nw, nb = p.nw, p.nb
    assert out.size == 2 * nw + nb
    return out[:nw], out[nw:2 * nw], out[2 * nw:]


def gppsum_cases():
    """Seeded problems for the gppsum = 1 loop: wtilde near the frequencies
    (so the unoccupied bands' near-pole terms are kept), ig ranges that
    cross the 512-wide blocks, columns at igp = 0, 1, 511, 512, 513, 1024
    (the diagonal term), the indigp == ncouls clamp, a skipped column, and
    occupied-band loops that stop early (|wtilde^2 I_eps| < to1)."""
    cases = {}
    cases["random"] = fko.gpp_random(6, 2, 1100, 11, nw=3, seed=41, wtilde_re=(2.5, 5.5))
    p = fko.gpp_random(5, 2, 1300, 12, nw=4, seed=42, wtilde_re=(2.5, 5.5))
    p.inv_igp_index = np.array([0, 1, 2, 511, 512, 513, 700, 1023, 1024, 1025, 1299, 1300], np.int32)
    ind = np.arange(1301, dtype=np.int32)
    ind[1300] = 1299
    ind[700] = -3            # skipped column
    p.indinv = ind
    # near-pole wtilde (|wx - wtilde| < 0.5 for wx = 3) at the points the
    # block bounds decide: the diagonal terms (ig = igp = 512, 1024), the
    # dropped ig = igp - 1 and ig = 511, and the first kept neighbours
    w = p.wtilde.copy()
    for m, ig in ((4, 512), (8, 1024), (4, 511), (8, 1023), (5, 511), (5, 512), (9, 1023), (9, 1022),
                  (10, 1298), (10, 1297), (7, 510), (7, 1022)):
        w[m, ig] = 3.1 + 0.02j
    p.wtilde = w
    cases["blocks"] = p
    q = fko.gpp_random(4, 3, 900, 9, nw=3, seed=43, wtilde_re=(2.5, 5.5))
    ie = q.I_eps.copy()
    for m, ig in ((2, 5), (4, 130), (6, 400), (7, 0), (8, 799)):
        ie[m, ig] = 1e-9 + 2e-9j     # |wtilde^2 I_eps| < to1: the occupied loop stops here
    q.I_eps = ie
    cases["break"] = q
    for name, p in cases.items():
        ach, asx, n1loc = run_gppseq_harness(p)
        np.savez_compressed(
            os.path.join(HERE, f"gppsum_{name}.npz"),
            nb=p.nb, nv=p.nv, ncouls=p.ncouls, ngpown=p.ngpown, occ=p.occ,
            wx=p.wx, aqsm=p.aqsm, aqsn=p.aqsn, I_eps=p.I_eps, wtilde=p.wtilde, vcoul=p.vcoul,
            inv_igp_index=p.inv_igp_index, indinv=p.indinv,
            achtemp=ach, asxtemp=asx, acht_n1_loc=n1loc)
        print(f"gppsum_{name}: achtemp={ach} asxtemp={asx}")
    shapes = [(8, 2, 1100, 10), (6, 3, 1536, 12), (4, 4, 600, 6)]
    rows = []
    for nb, nv, C, npg in shapes:
        out = subprocess.run([os.path.join(REF_BIN, "gppKerSeq"), str(nb), str(nv), str(C), str(npg), "1"],
                             check=True, capture_output=True, text=True).stdout
        vals = re.findall(r"achtemp\[(\d)\] = \(([-\d.e+]+),([-\d.e+]+)\)", out)
        assert len(vals) == 3, out
        rows.append([complex(float(a), float(b)) for _, a, b in vals])
    np.savez_compressed(os.path.join(HERE, "gppsum_main.npz"), shapes=np.array(shapes, np.int64),
                        achtemp=np.array(rows))
    print("gppsum_main:", rows)


def main():
    for exe in ("ref_gpp_harness", "ref_gppseq_harness", "ref_ff_harness", "gppKer", "gppKerSeq", "ffKernel"):
        if not os.path.exists(os.path.join(REF_BIN, exe)):
            sys.exit(f"missing oracle/_ref/{exe}: run `make -C oracle ref` where the reference is present")
    if sys.argv[1:] == ["gppsum"]:
        gppsum_cases()
        return
    if len(sys.argv) > 1:      # only the named ff_full cases
        ff_full_cases(only=set(sys.argv[1:]))
        return
    save_gpp("random", case_random())
    save_gpp("nearpole", case_nearpole())
    save_gpp("branch3", case_branch3())
    save_gpp("skipmap", case_skipmap())
    points()
    reference_main()
    ff_cases()
    ff_full_cases()
    gpp_achstemp_case()
    gppsum_cases()


if __name__ == "__main__":
    main()
