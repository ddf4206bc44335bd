"""ctypes wrapper of the CPU oracle (liboracle: oracle/libfko.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py as the checker; the product library never
imports, links or calls anything under oracle/.

Also holds the seeded synthetic-input builders shared by the tests: the
distributions are the ones the hot-path workloads are quoted on
(aqs*/I_eps re,im ~ U(-0.5,0.5); wtilde re ~ U(0.5,2.0), im ~ U(-0.01,0.01);
vcoul ~ U(0.1,10)), drawn from the counter hash in fko_rng.c so that the
device fill in the product library reproduces them bit for bit.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfko.so")

# array stream ids of the counter hash (shared with the device fill)
STREAM = {"aqsm": 1, "aqsn": 2, "I_eps": 3, "wtilde": 4, "vcoul": 5,
          "I_epsR": 6, "I_epsA": 7}

# reference constants, BGW-Kernels/gppKer.cpp:178-201
E_LK, DW, E_N1KQ = 10.0, 1.0, 6.0
TO1, LIMITTWO = 1e-6, 0.25
LIMITONE = 1.0 / (TO1 * 4.0)

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.fko_hash.restype = ctypes.c_uint64
        L.fko_hash.argtypes = [ctypes.c_uint64] * 3
        L.fko_uniform.restype = D
        L.fko_uniform.argtypes = [ctypes.c_uint64] * 3 + [D, D]
        L.fko_fill_real.argtypes = [P, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint64, D, D]
        L.fko_fill_complex.argtypes = [P, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint64, D, D, D, D]
        L.fko_gpp_point.restype = I
        L.fko_gpp_point.argtypes = [D, _Z, _Z, D, D, D, P, P]
        L.fko_gpp.restype = I
        L.fko_gpp.argtypes = [I, I, I, I, I] + [P] * 8 + [D, D, D, D] + [I] * 5 + [P] * 3
        L.fko_resolve_igp.restype = I
        L.fko_resolve_igp.argtypes = [I, I, P, P, P]
        L.fko_ff.restype = I
        L.fko_ff.argtypes = [I, I, I, I] + [P] * 6 + [I, I] + [P] * 3
        L.fko_gpp_gppsum.restype = I
        L.fko_gpp_gppsum.argtypes = [I, I, I, I, I] + [P] * 8 + [D, D, D, D] + [I] + [P] * 3
        L.fko_gppsum_unocc_kind.restype = I
        L.fko_gppsum_unocc_kind.argtypes = [I, I]
        L.fko_gpp_achstemp.restype = I
        L.fko_gpp_achstemp.argtypes = [I, I, I] + [P] * 6 + [I, P]
        L.fko_ff_full.restype = I
        L.fko_ff_full.argtypes = [ctypes.POINTER(_FfFullProblem), I, I, ctypes.POINTER(_FfFullOut)]
        _lib = L
    return _lib


class _Z(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


class _FfFullProblem(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_int) for n in ("nb", "nv", "ncouls", "ngpown", "nfreq", "nfreqeval")]
                + [(n, ctypes.c_void_p) for n in ("aqsm", "aqsn", "I_epsR", "I_epsA", "vcoul", "inv_igp_index",
                                                   "indinv", "ekq", "dFreqGrid", "pref", "dFreqBrd")]
                + [(n, ctypes.c_double) for n in ("pref_zb", "freqevalmin", "freqevalstep", "occ")]
                + [("achs_freq", ctypes.c_int), ("schDt_in", ctypes.c_void_p)])


class _FfFullOut(ctypes.Structure):
    _fields_ = [("achsDtemp", _Z)] + [(n, ctypes.c_void_p) for n in (
        "asxDtemp", "achDtemp", "achDtemp_cor", "achDtemp_corb", "ach2Dtemp", "schDt")]


def _p(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


# --------------------------------------------------------------------------
# seeded inputs
# --------------------------------------------------------------------------
def fill_real(n: int, seed: int, stream: int, lo: float, hi: float) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().fko_fill_real(_p(out), n, seed, stream, lo, hi)
    return out


def fill_complex(n: int, seed: int, stream: int, re: tuple, im: tuple) -> np.ndarray:
    out = np.empty(n, dtype=np.complex128)
    lib().fko_fill_complex(_p(out), n, seed, stream, re[0], re[1], im[0], im[1])
    return out


def wx_array(nw: int) -> np.ndarray:
    """wx_array[iw] = e_lk - e_n1kq + dw*(iw-1), clamped (BGW-Kernels/gppKer.cpp:310-314)."""
    wx = np.array([E_LK - E_N1KQ + DW * ((iw + 1) - 2) for iw in range(nw)], dtype=np.float64)
    wx[np.abs(wx) < TO1] = TO1
    return wx


def index_maps(ncouls: int, ngpown: int, kind: str = "fortran"):
    """Column maps.  'cpp': inv_igp_index[i] = (i+1)*C/G (BGW-Kernels/gppKer.cpp:294-295);
    'fortran': i*C/G (gppkernel.deslippe.09.30.2014.f90:150-152, 0-based).
    indinv[i] = i with the extra entry indinv[C] = C-1 (:298-299)."""
    i = np.arange(ngpown, dtype=np.int64)
    if kind == "cpp":
        inv = (i + 1) * ncouls // ngpown
    elif kind == "fortran":
        inv = i * ncouls // ngpown
    else:
        raise ValueError(kind)
    indinv = np.arange(ncouls + 1, dtype=np.int32)
    indinv[ncouls] = ncouls - 1
    return inv.astype(np.int32), indinv


@dataclass
class GppProblem:
    nb: int
    nv: int
    ncouls: int
    ngpown: int
    wx: np.ndarray
    aqsm: np.ndarray      # [nb][C] complex128
    aqsn: np.ndarray      # [nb][C]
    I_eps: np.ndarray     # [G][C]
    wtilde: np.ndarray    # [G][C]
    vcoul: np.ndarray     # [C]
    inv_igp_index: np.ndarray   # [G] int32
    indinv: np.ndarray          # [C+1] int32
    to1: float = TO1
    limittwo: float = LIMITTWO
    limitone: float = LIMITONE
    occ: float = 1.0
    meta: dict = field(default_factory=dict)

    @property
    def nw(self) -> int:
        return int(self.wx.shape[0])

    @property
    def updates(self) -> int:
        return self.nb * self.ngpown * self.ncouls * self.nw


def gpp_random(nb, nv, ncouls, ngpown, nw=3, seed=0, wtilde_re=(0.5, 2.0),
               map_kind="fortran") -> GppProblem:
    """Seeded inputs with the distributions of the quoted GPP workloads."""
    C, G = ncouls, ngpown
    aqsm = fill_complex(nb * C, seed, STREAM["aqsm"], (-0.5, 0.5), (-0.5, 0.5)).reshape(nb, C)
    aqsn = fill_complex(nb * C, seed, STREAM["aqsn"], (-0.5, 0.5), (-0.5, 0.5)).reshape(nb, C)
    I_eps = fill_complex(G * C, seed, STREAM["I_eps"], (-0.5, 0.5), (-0.5, 0.5)).reshape(G, C)
    wtilde = fill_complex(G * C, seed, STREAM["wtilde"], wtilde_re, (-0.01, 0.01)).reshape(G, C)
    vcoul = fill_real(C, seed, STREAM["vcoul"], 0.1, 10.0)
    inv, indinv = index_maps(C, G, map_kind)
    return GppProblem(nb, nv, C, G, wx_array(nw), aqsm, aqsn, I_eps, wtilde, vcoul, inv, indinv,
                      meta={"seed": seed, "wtilde_re": wtilde_re, "map": map_kind})


def gpp_constant(nb, nv, ncouls, nodes_per_group, map_kind="cpp") -> GppProblem:
    """The reference driver's own inputs: every complex entry 0.5+0.5i, vcoul = 1
    (BGW-Kernels/gppKer.cpp:275-299), ngpown = ncouls / nodes_per_group (:176)."""
    C = ncouls
    G = ncouls // nodes_per_group
    z = np.full((nb, C), 0.5 + 0.5j)
    w = np.full((G, C), 0.5 + 0.5j)
    inv, indinv = index_maps(C, G, map_kind)
    return GppProblem(nb, nv, C, G, wx_array(3), z, z.copy(), w, w.copy(), np.ones(C), inv, indinv,
                      meta={"constant": True, "map": map_kind})


# --------------------------------------------------------------------------
# oracle calls
# --------------------------------------------------------------------------
def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def gpp(p: GppProblem, mask_on=True, extended=True, nthreads=0, band_lo=0, band_hi=None,
        want_n1loc=False):
    """Run the oracle band loop.  Returns (achtemp[nw], asxtemp[nw], acht_n1_loc[nb] or None)."""
    band_hi = p.nb if band_hi is None else band_hi
    ach = np.zeros(p.nw, np.complex128)
    asx = np.zeros(p.nw, np.complex128)
    n1loc = np.zeros(max(p.nb, 1), np.complex128) if want_n1loc else None
    arrs = [_c(p.wx, np.float64), _c(p.aqsm, np.complex128), _c(p.aqsn, np.complex128),
            _c(p.I_eps, np.complex128), _c(p.wtilde, np.complex128), _c(p.vcoul, np.float64),
            _c(p.inv_igp_index, np.int32), _c(p.indinv, np.int32)]
    rc = lib().fko_gpp(p.nb, p.nv, p.ncouls, p.ngpown, p.nw, *[_p(a) for a in arrs],
                       p.to1, p.limittwo, p.limitone, p.occ,
                       int(mask_on), int(extended), int(nthreads), int(band_lo), int(band_hi),
                       _p(ach), _p(asx), _p(n1loc) if want_n1loc else None)
    if rc != 0:
        raise ValueError("oracle rejected the GPP problem")
    return ach, asx, (n1loc[: p.nb] if want_n1loc else None)


def gpp_gppsum(p: GppProblem, nthreads=0, want_n1loc=False):
    """The gppsum = 1 band loop of gppKerSeq.cpp.  Returns (achtemp[nw],
    asxtemp[nw], acht_n1_loc[nb] or None)."""
    ach = np.zeros(p.nw, np.complex128)
    asx = np.zeros(p.nw, np.complex128)
    n1loc = np.zeros(max(p.nb, 1), np.complex128) if want_n1loc else None
    arrs = [_c(p.wx, np.float64), _c(p.aqsm, np.complex128), _c(p.aqsn, np.complex128),
            _c(p.I_eps, np.complex128), _c(p.wtilde, np.complex128), _c(p.vcoul, np.float64),
            _c(p.inv_igp_index, np.int32), _c(p.indinv, np.int32)]
    rc = lib().fko_gpp_gppsum(p.nb, p.nv, p.ncouls, p.ngpown, p.nw, *[_p(a) for a in arrs],
                              p.to1, p.limittwo, p.limitone, p.occ, int(nthreads),
                              _p(ach), _p(asx), _p(n1loc) if want_n1loc else None)
    if rc != 0:
        raise ValueError("oracle rejected the GPP problem")
    return ach, asx, (n1loc[: p.nb] if want_n1loc else None)


def gppsum_unocc_kind(ig, igp):
    return int(lib().fko_gppsum_unocc_kind(int(ig), int(igp)))


def gpp_point(wxt, wtilde, I_eps, to1=TO1, limittwo=LIMITTWO, limitone=LIMITONE):
    ssx = np.zeros(1, np.complex128)
    sch = np.zeros(1, np.complex128)
    br = lib().fko_gpp_point(wxt, _Z(wtilde.real, wtilde.imag), _Z(I_eps.real, I_eps.imag),
                             to1, limittwo, limitone, _p(ssx), _p(sch))
    return br, complex(ssx[0]), complex(sch[0])


def gpp_achstemp(p, nthreads=0):
    """Static-CH remainder of the Fortran GPP program over every band
    (gppkernel.deslippe.09.30.2014.f90:192-230) for a GppProblem."""
    out = np.zeros(1, np.complex128)
    arrs = [_c(p.aqsm, np.complex128), _c(p.aqsn, np.complex128), _c(p.I_eps, np.complex128),
            _c(p.vcoul, np.float64), _c(p.inv_igp_index, np.int32), _c(p.indinv, np.int32)]
    rc = lib().fko_gpp_achstemp(p.nb, p.ncouls, p.ngpown, *[_p(a) for a in arrs], int(nthreads), _p(out))
    if rc != 0:
        raise ValueError("oracle rejected the problem")
    return complex(out[0])


def resolve_igp(ncouls, ngpown, inv, indinv):
    out = np.zeros(ngpown, np.int32)
    rc = lib().fko_resolve_igp(ncouls, ngpown, _p(_c(inv, np.int32)), _p(_c(indinv, np.int32)), _p(out))
    if rc != 0:
        raise ValueError("index map out of range")
    return out


@dataclass
class FfProblem:
    nb: int
    ncouls: int
    ngpown: int
    nfreq: int
    aqsm: np.ndarray     # [nb][C]
    aqsn: np.ndarray     # [nb][C]
    I_eps: np.ndarray    # [nfreq][G][C]
    vcoul: np.ndarray    # [C]
    inv_igp_index: np.ndarray
    indinv: np.ndarray   # [C+1]
    meta: dict = field(default_factory=dict)


def ff_random(nb, ncouls, ngpown, nfreq, seed=0, which="I_epsR") -> FfProblem:
    C, G = ncouls, ngpown
    aqsm = fill_complex(nb * C, seed, STREAM["aqsm"], (-0.5, 0.5), (-0.5, 0.5)).reshape(nb, C)
    aqsn = fill_complex(nb * C, seed, STREAM["aqsn"], (-0.5, 0.5), (-0.5, 0.5)).reshape(nb, C)
    I_eps = fill_complex(nfreq * G * C, seed, STREAM[which], (-0.5, 0.5), (-0.5, 0.5)).reshape(nfreq, G, C)
    vcoul = fill_real(C, seed, STREAM["vcoul"], 0.1, 10.0)
    inv = np.arange(G, dtype=np.int32)          # FFKer_cori/ffKernel.cpp:404-405
    indinv = np.arange(C + 1, dtype=np.int32)
    indinv[C] = C - 1
    return FfProblem(nb, C, G, nfreq, aqsm, aqsn, I_eps, vcoul, inv, indinv, meta={"seed": seed})


def ff(p: FfProblem, achs_freq=0, nthreads=0):
    """Returns (schDt[nb][nfreq], schDv[nb][nfreq], achsDtemp)."""
    schDt = np.zeros((p.nb, p.nfreq), np.complex128)
    schDv = np.zeros((p.nb, p.nfreq), np.complex128)
    achs = np.zeros(1, np.complex128)
    arrs = [_c(p.aqsm, np.complex128), _c(p.aqsn, np.complex128), _c(p.I_eps, np.complex128),
            _c(p.vcoul, np.float64), _c(p.inv_igp_index, np.int32), _c(p.indinv, np.int32)]
    rc = lib().fko_ff(p.nb, p.ncouls, p.ngpown, p.nfreq, *[_p(a) for a in arrs],
                      int(achs_freq), int(nthreads), _p(schDt), _p(schDv), _p(achs))
    if rc != 0:
        raise ValueError("oracle rejected the FF problem")
    return schDt, schDv, complex(achs[0])


# --------------------------------------------------------------------------
# closed forms of the constant-input case (derived from the formulas of
# BGW-Kernels/gppKer.cpp:19-65,394-411 with every entry 0.5+0.5i, vcoul=1)
# --------------------------------------------------------------------------
def gpp_closed_form(nb, nv, ncouls, ngpown, wx):
    """achtemp[iw] = nb*G*C*0.25*(-0.25 + i(0.5 wx - 0.25)) / ((wx-0.5)^2 + 0.25)
    asxtemp[iw] = nv*G*C*(-0.125 + 0.125i)/wx^2   (mask branch everywhere)."""
    wx = np.asarray(wx, dtype=np.float64)
    n = float(nb) * ngpown * ncouls
    ach = n * 0.25 * (-0.25 + 1j * (0.5 * wx - 0.25)) / ((wx - 0.5) ** 2 + 0.25)
    asx = float(nv) * ngpown * ncouls * (-0.125 + 0.125j) / wx ** 2
    return ach, asx


# --------------------------------------------------------------------------
# whole FF evaluation (fko_ff_full.c)
# --------------------------------------------------------------------------
FF_OUTPUTS = ("asxDtemp", "achDtemp", "achDtemp_cor", "achDtemp_corb", "ach2Dtemp")


@dataclass
class FfFullProblem:
    nb: int
    nv: int
    ncouls: int
    ngpown: int
    nfreq: int
    nfreqeval: int
    aqsm: np.ndarray      # [nb][C]
    aqsn: np.ndarray      # [nb][C]
    I_epsR: np.ndarray    # [nfreq][G][C]
    I_epsA: np.ndarray    # [nfreq][G][C]
    vcoul: np.ndarray     # [C]
    inv_igp_index: np.ndarray
    indinv: np.ndarray    # [C+1]
    ekq: np.ndarray       # [nb]
    dFreqGrid: np.ndarray  # [nfreq]
    pref: np.ndarray      # [nfreq]
    dFreqBrd: np.ndarray  # [nfreq] complex
    pref_zb: float = 0.5 / 3.14
    freqevalmin: float = 0.0
    freqevalstep: float = 0.5
    occ: float = 1.0
    achs_freq: int = 0
    meta: dict = field(default_factory=dict)


def ff_program_axes(nb, nfreq):
    """The FF program's band energies and frequency grid
    (ffkernel.deslippe.best.f90:124-171): ekq[n1] = -10 + n1,
    dFreqGrid[f] = 2 f, pref[f] = (dF[f+1] - dF[f]) / 3.14 (last = previous;
    first and last halved), dFreqBrd = 0.1i."""
    ekq = -10.0 + np.arange(nb, dtype=np.float64)
    grid = 2.0 * np.arange(nfreq, dtype=np.float64)
    pref = np.empty(nfreq)
    for f in range(nfreq):
        pref[f] = (grid[f + 1] - grid[f]) / 3.14 if f < nfreq - 1 else pref[f - 1]
    pref[0] *= 0.5
    pref[nfreq - 1] *= 0.5
    brd = np.full(nfreq, 0.1j, dtype=np.complex128)
    return ekq, grid, pref, brd


def ff_full_random(nb, nv, ncouls, ngpown, nfreq, nfreqeval, seed=0) -> FfFullProblem:
    """Seeded inputs with the program's axes: aqs*, I_epsR, I_epsA ~
    U(-0.5,0.5) + iU(-0.5,0.5), vcoul ~ U(0.1,10), identity column maps."""
    base = ff_random(nb, ncouls, ngpown, nfreq, seed=seed, which="I_epsR")
    I_epsA = fill_complex(nfreq * ngpown * ncouls, seed, STREAM["I_epsA"], (-0.5, 0.5),
                          (-0.5, 0.5)).reshape(nfreq, ngpown, ncouls)
    ekq, grid, pref, brd = ff_program_axes(nb, nfreq)
    return FfFullProblem(nb, nv, ncouls, ngpown, nfreq, nfreqeval, base.aqsm, base.aqsn, base.I_eps, I_epsA,
                         base.vcoul, base.inv_igp_index, base.indinv, ekq, grid, pref, brd, meta={"seed": seed})


def ff_full(p: FfFullProblem, variant=0, schDt_in=None, nthreads=0):
    """Returns a dict: achsDtemp (complex), asxDtemp, achDtemp, achDtemp_cor,
    achDtemp_corb, ach2Dtemp ([nfreqeval] complex), schDt ([nb][nfreq])."""
    keep = [_c(p.aqsm, np.complex128), _c(p.aqsn, np.complex128), _c(p.I_epsR, np.complex128),
            _c(p.I_epsA, np.complex128), _c(p.vcoul, np.float64), _c(p.inv_igp_index, np.int32),
            _c(p.indinv, np.int32), _c(p.ekq, np.float64), _c(p.dFreqGrid, np.float64), _c(p.pref, np.float64),
            _c(p.dFreqBrd, np.complex128)]
    sin = None if schDt_in is None else _c(schDt_in, np.complex128)
    pr = _FfFullProblem(p.nb, p.nv, p.ncouls, p.ngpown, p.nfreq, p.nfreqeval, *[a.ctypes.data for a in keep],
                        p.pref_zb, p.freqevalmin, p.freqevalstep, p.occ, int(p.achs_freq),
                        None if sin is None else sin.ctypes.data)
    outs = {k: np.zeros(p.nfreqeval, np.complex128) for k in FF_OUTPUTS}
    outs["schDt"] = np.zeros((p.nb, p.nfreq), np.complex128)
    o = _FfFullOut(_Z(0, 0), *[outs[k].ctypes.data for k in FF_OUTPUTS + ("schDt",)])
    rc = lib().fko_ff_full(ctypes.byref(pr), int(variant), int(nthreads), ctypes.byref(o))
    if rc != 0:
        raise ValueError("oracle rejected the FF problem")
    outs["achsDtemp"] = complex(o.achsDtemp.re, o.achsDtemp.im)
    return outs
