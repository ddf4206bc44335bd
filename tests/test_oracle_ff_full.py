"""Pin the oracle's whole FF evaluation (oracle/fko_ff_full.c) before it is
used as the checker of ff_kernel_full.

Variant 1 (the C++ port FFKer_cori/ffKernel.cpp, one thread) is compared with
golden vectors its compiled kernel functions wrote on seeded inputs, and with
the values the compiled ffKernel program printed.  It shares the
interpolated SX / delta contractions and the sigma-1 arithmetic with
variant 0 (the Fortran program, canonical), which is then pinned by:
  * closed forms on inputs constant over (my_igp, ig) and linear or
    quadratic in the grid frequency: the SX and delta interpolation is exact
    on a linear I_eps and gives the chord of the interval it picks on a
    quadratic one; the sigma-2 / sigma-4 log sums telescope;
  * the tie of variant 0's sigma 1 to variant 1's (evaluation axis shifted by
    one step, last weight zero);
  * symmetries: sigma 3 equals sigma 1 for unoccupied bands, and for occupied
    bands it is the conjugate of sigma 1 at -wx with conjugated inputs.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

OUTS = ("asxDtemp", "achDtemp", "ach2Dtemp", "achDtemp_cor", "achDtemp_corb")


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _load(fko, name):
    d = np.load(os.path.join(GOLDEN, f"ff_full_{name}.npz"))
    nb, nv, C, G, nF, ne = (int(x) for x in d["sizes"])
    pz, fmin, fstep, occ = (float(x) for x in d["scalars"])
    p = fko.FfFullProblem(nb, nv, C, G, nF, ne, d["aqsm"], d["aqsn"], d["I_epsR"], d["I_epsA"], d["vcoul"],
                          d["inv_igp_index"], d["indinv"], d["ekq"], d["dFreqGrid"], d["pref"], d["dFreqBrd"],
                          pz, fmin, fstep, occ, achs_freq=1)
    return p, d


@pytest.mark.parametrize("name", ["seeded", "program_axes", "negative_axes"])
def test_cpp_variant_matches_reference_functions(fko, name):
    p, d = _load(fko, name)
    o = fko.ff_full(p, variant=1, schDt_in=d["schDt"])
    for k in OUTS:
        ref = d[k]
        if np.max(np.abs(ref)) == 0:
            assert np.max(np.abs(o[k])) == 0, k
        else:
            assert _rel(o[k], ref) <= 1e-13, (k, _rel(o[k], ref))
    assert abs(o["achsDtemp"] - complex(d["achsDtemp"])) <= 1e-13 * abs(complex(d["achsDtemp"]))


def _constant_problem(fko, nb, nv, C, G, nF, ne, q=0.5 + 0.5j, R=(0.5 + 0.5j, 0.0), A=(0.5 - 0.5j, 0.0),
                      v=1.0, ekq_shift=0.0, quad=False):
    """aqsm = aqsn = q, vcoul = v, I_epsX[f] = a + b g(dF[f]) with g(x) = x
    (or x^2 with quad=True), identity maps, the program's axes."""
    ekq, grid, pref, brd = fko.ff_program_axes(nb, nF)
    g = grid ** 2 if quad else grid
    IR = np.broadcast_to((R[0] + R[1] * g)[:, None, None], (nF, G, C)).copy()
    IA = np.broadcast_to((A[0] + A[1] * g)[:, None, None], (nF, G, C)).copy()
    inv = np.arange(G, dtype=np.int32)
    indinv = np.arange(C + 1, dtype=np.int32)
    indinv[C] = C - 1
    z = np.full((nb, C), q, np.complex128)
    return fko.FfFullProblem(nb, nv, C, G, nF, ne, z, z.copy(), IR, IA, np.full(C, v), inv, indinv,
                             ekq + ekq_shift, grid, pref, brd)


def _wx(p):
    iw = np.arange(p.nfreqeval)
    return p.freqevalmin - p.ekq[:, None] + iw[None, :] * p.freqevalstep      # [nb][ne]


def test_program_output_reproduced_by_cpp_variant(fko):
    """The ffKernel program's printed asxDtemp[0] and achDtemp_cor[0]
    (constant inputs, schDt_matrix left zero by its main; printed with %f)."""
    d = np.load(os.path.join(GOLDEN, "ff_main.npz"))
    for (nb, nv, C, G, nF, ne), asx, cor in zip(d["shapes"], d["asxDtemp"], d["achDtemp_cor"]):
        p = _constant_problem(fko, int(nb), int(nv), int(C), int(G), int(nF), int(ne))
        o = fko.ff_full(p, variant=1)
        assert abs(o["asxDtemp"][0] - asx) <= 1e-6 * max(1.0, abs(asx))
        assert abs(o["achDtemp_cor"][0] - cor) <= 1e-6 * max(1.0, abs(cor))


@pytest.mark.parametrize("fmin", [0.0, 0.5])
def test_closed_form_linear_inputs(fko, fmin):
    """Linear-in-frequency I_eps: SX and delta interpolation are exact; the
    sigma-4 sums telescope to S(w) L + S(dF_last) - S(dF_0)."""
    nb, nv, C, G, nF, ne = 24, 9, 6, 4, 7, 10
    q, v = 0.3 - 0.7j, 1.7
    R, A = (0.4 + 0.2j, 0.05 - 0.03j), (-0.1 + 0.3j, -0.02 + 0.04j)
    p = _constant_problem(fko, nb, nv, C, G, nF, ne, q=q, R=R, A=A, v=v, ekq_shift=-0.25)
    p.freqevalmin = fmin
    o = fko.ff_full(p)
    K = G * C * abs(q) ** 2                  # schDt = K I_R(.), schDv = v K I(.)
    wx = _wx(p)
    occ = np.arange(nb) < nv
    dF = p.dFreqGrid
    S = lambda x: K * (R[0] + R[1] * x)      # noqa: E731
    dRA = lambda x: (R[0] - A[0]) + (R[1] - A[1]) * x   # noqa: E731
    # SX: interpolated at |wx|, I_epsR for wx >= 0, I_epsA below
    sx = np.where(wx >= 0, R[0] + R[1] * np.abs(wx), A[0] + A[1] * np.abs(wx)) * v * K
    assert _rel(o["asxDtemp"], (sx * occ[:, None]).sum(0) * p.occ) < 1e-13
    # delta terms
    d_pos = -0.5 * v * K * dRA(wx)
    d_neg = -0.5 * v * K * dRA(-wx)
    # sigma 2 with a linear S: pairwise average times the log ratio (no clamp
    # is reached on these axes: |wx - dF| >= 0.25)
    zr, zl = dF[1:], dF[:-1]
    it = -np.log(np.abs((wx[..., None] - zr) / (wx[..., None] - zl)))
    assert np.all(np.abs((wx[..., None] - zr) / (wx[..., None] - zl)) > 1e-4)
    avg = 0.5 * (S(zr) + S(zl))
    s2 = (-1j * p.pref_zb * avg * it).sum(-1)
    a2 = (s2 + np.where(wx >= 0, d_pos, 0)).sum(0)
    assert _rel(o["ach2Dtemp"], a2) < 1e-13
    # sigma 4, telescoped
    L_un = np.log(np.abs(wx - dF[0])) - np.log(np.abs(wx - dF[-1]))
    L_oc = np.log(np.abs(wx + dF[-1])) - np.log(np.abs(wx + dF[0]))
    wxp = -p.freqevalmin - p.ekq[:, None] + np.arange(ne)[None, :] * p.freqevalstep
    s4 = np.where(occ[:, None], S(wxp) * L_oc, S(wx) * L_un) + (S(dF[-1]) - S(dF[0]))
    cor = -1j * p.pref_zb * s4
    cor = cor + np.where(wx >= 0, np.where(occ[:, None], 0, d_pos), np.where(occ[:, None], d_neg, 0))
    assert _rel(o["achDtemp_cor"], cor.sum(0)) < 1e-12
    # static CH at the first frequency: -0.5 v K I_R(dF_0) per band
    assert abs(o["achsDtemp"] - nb * (-0.5) * v * K * (R[0] + R[1] * dF[0])) < 1e-13 * abs(o["achsDtemp"])


@pytest.mark.parametrize("shift", [-0.25, 0.0])
def test_interval_choice_on_quadratic_inputs(fko, shift):
    """I_eps = a + b dF^2: linear interpolation gives the chord of the chosen
    interval, (dF_i + dF_i+1) w - dF_i dF_i+1, so the interval index matters.
    Grid 2f: interval floor(w / 2), or the last one (nFreq-2) above the grid.
    shift 0 puts wx on grid points, 0 included (dF_0 <= w < dF_1)."""
    nb, nv, C, G, nF, ne = 40, 40, 3, 2, 6, 10
    R, A = (0.2 + 0.1j, 0.01 - 0.02j), (0.3 - 0.1j, -0.015 + 0.01j)
    p = _constant_problem(fko, nb, nv, C, G, nF, ne, q=1.0, R=R, A=A, v=1.0, ekq_shift=shift, quad=True)
    o = fko.ff_full(p)
    wx = _wx(p)
    K = G * C
    w = np.abs(wx)
    i = np.minimum(np.floor(w / 2.0).astype(int), nF - 2)
    lo, hi = 2.0 * i, 2.0 * (i + 1)
    chord = (lo + hi) * w - lo * hi
    sx = np.where(wx >= 0, R[0] + R[1] * chord, A[0] + A[1] * chord) * K
    assert _rel(o["asxDtemp"], sx.sum(0)) < 1e-13
    dchord = (R[0] - A[0]) + (R[1] - A[1]) * chord
    # all bands occupied: wx >= 0 -> sch2Di only; wx < 0 -> achDtemp_cor only
    delta_cor = np.where(wx < 0, -0.5 * K * dchord, 0).sum(0)
    p_sig = _constant_problem(fko, nb, nv, C, G, nF, ne, q=1.0, R=R, A=R, v=1.0, ekq_shift=shift, quad=True)
    o_sig = fko.ff_full(p_sig)              # I_epsA == I_epsR: no delta terms
    assert _rel(o["achDtemp_cor"] - o_sig["achDtemp_cor"], delta_cor) < 1e-12
    delta_2 = np.where(wx >= 0, -0.5 * K * dchord, 0).sum(0)
    assert _rel(o["ach2Dtemp"] - o_sig["ach2Dtemp"], delta_2) < 1e-12


def test_sigma1_tied_to_cpp_variant(fko):
    """Variant 1 computes sigma 1 over every frequency at w - step into
    asxDtemp; with no occupied band (no SX term), the last weight zero and the
    evaluation axis moved down one step, it is variant 0's achDtemp."""
    p = fko.ff_full_random(10, 0, 20, 5, 6, 8, seed=3)
    p.pref = p.pref.copy()
    p.pref[-1] = 0.0
    o1 = fko.ff_full(p, variant=1, schDt_in=fko.ff_full(p)["schDt"])
    p.freqevalmin -= p.freqevalstep
    o0 = fko.ff_full(p)
    assert _rel(o0["achDtemp"], o1["asxDtemp"]) < 1e-13


def test_frequency_ranges_of_the_sigma_terms(fko):
    """Sigma 1 / 3 run over every frequency but the last, sigma 2 / 4 over
    every frequency but the first: a schDt_matrix that is zero except in the
    last column leaves achDtemp and achDtemp_corb zero, and one that is zero
    except in the first column leaves ach2Dtemp and achDtemp_cor equal to the
    delta terms alone (those come from I_epsR - I_epsA, not from schDt)."""
    p = fko.ff_full_random(8, 3, 12, 4, 5, 6, seed=13)
    zero = fko.ff_full(p, schDt_in=np.zeros((p.nb, p.nfreq)))
    last = np.zeros((p.nb, p.nfreq), complex)
    last[:, -1] = 1.0 + 2.0j
    o = fko.ff_full(p, schDt_in=last)
    assert np.max(np.abs(o["achDtemp"])) == 0 and np.max(np.abs(o["achDtemp_corb"])) == 0
    assert np.max(np.abs(o["ach2Dtemp"] - zero["ach2Dtemp"])) > 0
    first = np.zeros((p.nb, p.nfreq), complex)
    first[:, 0] = 1.0 + 2.0j
    o = fko.ff_full(p, schDt_in=first)
    assert np.max(np.abs(o["achDtemp"])) > 0
    # the first column still enters sigma 2 / 4 as schDt_left of frequency 1
    assert np.max(np.abs(o["ach2Dtemp"] - zero["ach2Dtemp"])) > 0


def test_sigma3_symmetries(fko):
    p = fko.ff_full_random(9, 0, 20, 5, 6, 8, seed=5)
    o = fko.ff_full(p)
    assert _rel(o["achDtemp_corb"], o["achDtemp"]) == 0.0       # unoccupied: same formula
    # occupied bands: sigma3(w; S) = conj(sigma1(-w; conj S)); -w via
    # freqevalmin, ekq, freqevalstep -> their negatives
    p.nv = p.nb
    o_occ = fko.ff_full(p)
    pc = fko.ff_full_random(9, 0, 20, 5, 6, 8, seed=5)
    pc.aqsm, pc.aqsn, pc.I_epsR = np.conj(pc.aqsm), np.conj(pc.aqsn), np.conj(pc.I_epsR)
    pc.ekq, pc.freqevalmin, pc.freqevalstep = -pc.ekq, -pc.freqevalmin, -pc.freqevalstep
    o_c = fko.ff_full(pc)
    assert _rel(o_occ["achDtemp_corb"], np.conj(o_c["achDtemp"])) < 1e-13


def test_sigma2_constant_telescopes(fko):
    """Constant S: the sigma-2 sum is -i pref_zb S (log|w - dF_0| - log|w - dF_last|)."""
    nb, nv, C, G, nF, ne = 16, 0, 4, 3, 8, 6
    p = _constant_problem(fko, nb, nv, C, G, nF, ne, ekq_shift=-0.25, A=(0.5 + 0.5j, 0.0))
    o = fko.ff_full(p)                      # I_epsA == I_epsR: no delta terms
    S = G * C * 0.5 * (0.5 + 0.5j)
    wx = _wx(p)
    L = np.log(np.abs(wx - p.dFreqGrid[0])) - np.log(np.abs(wx - p.dFreqGrid[-1]))
    assert _rel(o["ach2Dtemp"], (-1j * p.pref_zb * S * L).sum(0)) < 1e-13


def test_variants_differ_where_the_port_departs(fko):
    p = fko.ff_full_random(12, 3, 24, 6, 5, 9, seed=42)
    o0, o1 = fko.ff_full(p), fko.ff_full(p, variant=1, schDt_in=fko.ff_full(p)["schDt"])
    assert np.max(np.abs(o1["achDtemp"])) == 0 and np.max(np.abs(o0["achDtemp"])) > 0
    assert _rel(o0["asxDtemp"], o1["asxDtemp"]) > 1e-3


def test_thread_count_independent(fko):
    p = fko.ff_full_random(11, 4, 30, 7, 5, 6, seed=8)
    a, b = fko.ff_full(p, nthreads=1), fko.ff_full(p, nthreads=3)
    for k in OUTS + ("schDt",):
        assert np.array_equal(a[k], b[k]), k
