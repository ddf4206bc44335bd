"""GPU parity of ff_kernel_full (the whole FF evaluation of the Fortran
program ffkernel.deslippe.best.f90: static CH, SX, sigma 1-4, delta terms)
against the oracle's canonical variant on the same complex128 inputs.

FP64: every output within 1e-12 (normwise, max |diff| / max |ref| over the
evaluation frequencies) of the oracle, which restates the Fortran with direct
sums; the device contracts first and interpolates after.  At 256/8/4096/4096/
32/10 the program's constant inputs make every output proportional to
ngpown * ncouls, which the oracle evaluates at small ncouls / ngpown.
Mixed precision (both contractions in split BF16): the error against the
oracle is stated and bounded by MIXED_RTOL; measured on MI355X 5e-8 .. 6e-7
on the small cases and up to 9.4e-7 at full size against the FP64 path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-12
MIXED_RTOL = 5e-6
OUTS = ("asxDtemp", "achDtemp", "achDtemp_cor", "achDtemp_corb", "ach2Dtemp")


def _torch():
    import torch
    assert torch.cuda.is_available(), "no ROCm device visible"
    return torch


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _to_dev(p):
    torch = _torch()
    d = torch.device("cuda:0")
    f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(d).to(dt)  # noqa: E731
    q = type(p)(**{k: getattr(p, k) for k in p.__dataclass_fields__})
    q.aqsm, q.aqsn = f(p.aqsm, torch.complex128), f(p.aqsn, torch.complex128)
    q.I_epsR, q.I_epsA = f(p.I_epsR, torch.complex128), f(p.I_epsA, torch.complex128)
    q.vcoul = f(p.vcoul, torch.float64)
    return q


def _np(res):
    out = {}
    for k, v in res.items():
        if v is None or isinstance(v, float):
            out[k] = v
        else:
            out[k] = v.cpu().numpy() if hasattr(v, "cpu") else v
    out["achsDtemp"] = complex(np.asarray(out["achsDtemp"]).reshape(-1)[0])
    return out


def _check(got, ref, tol, label):
    errs = {k: _rel(got[k], ref[k]) for k in OUTS}
    errs["achsDtemp"] = abs(got["achsDtemp"] - ref["achsDtemp"]) / abs(ref["achsDtemp"])
    errs["schDt"] = _rel(got["schDt"], ref["schDt"])
    print(f"  {label}: " + " ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v <= tol, (label, k, v)


def _variety(fko, seed, nb, nv, C, G, nF, ne, fmin=0.0, step=0.5, random_axes=False):
    p = fko.ff_full_random(nb, nv, C, G, nF, ne, seed=seed)
    p.freqevalmin, p.freqevalstep = fmin, step
    if random_axes:
        rng = np.random.default_rng(seed)
        p.ekq = rng.uniform(-1.5 * nF, 1.5 * nF, nb)
        p.dFreqGrid = np.cumsum(rng.uniform(0.3, 3.0, nF))
        p.dFreqGrid -= p.dFreqGrid[0]
        p.pref = rng.uniform(0.05, 0.7, nF)
        p.dFreqBrd = rng.uniform(-0.05, 0.05, nF) + 1j * rng.uniform(0.05, 0.3, nF)
        p.occ = 0.5
        p.pref_zb = 0.2
    return p


CASES = [
    # seed, nb, nv, C, G, nF, ne, fmin, step, random axes
    (1, 12, 3, 40, 8, 6, 10, 0.0, 0.5, False),
    (2, 30, 8, 100, 30, 8, 7, 0.3, 0.7, False),
    (3, 70, 20, 260, 64, 12, 5, -1.0, 1.25, True),
    (4, 300, 9, 96, 40, 5, 11, 0.25, 0.5, True),
    (5, 9, 9, 33, 33, 2, 3, 0.0, 0.5, False),
    (6, 5, 0, 17, 3, 4, 1, 0.0, 0.5, True),
]


@pytest.mark.parametrize("case", CASES, ids=[f"s{c[0]}" for c in CASES])
def test_fp64_vs_oracle(fk, fko, case):
    seed, nb, nv, C, G, nF, ne, fmin, step, ra = case
    print(f"ff_full fp64 case {case}", flush=True)
    p = _variety(fko, seed, nb, nv, C, G, nF, ne, fmin, step, ra)
    ref = fko.ff_full(p)
    got = _np(fk.ff_full_problem(_to_dev(p)))
    _check(got, ref, RTOL, f"nb={nb} nv={nv} C={C} G={G} nF={nF} ne={ne}")


@pytest.mark.parametrize("case", CASES[:4], ids=[f"s{c[0]}" for c in CASES[:4]])
def test_mixed_vs_oracle(fk, fko, case):
    seed, nb, nv, C, G, nF, ne, fmin, step, ra = case
    print(f"ff_full mixed case {case}", flush=True)
    p = _variety(fko, seed, nb, nv, C, G, nF, ne, fmin, step, ra)
    ref = fko.ff_full(p)
    got = _np(fk.ff_full_problem(_to_dev(p), precision="mixed"))
    _check(got, ref, MIXED_RTOL, "mixed")


def test_no_band_reads_I_epsA(fk, fko):
    """No occupied band and every wx < 0: neither the SX nor the delta terms
    read I_epsA, so its contraction is skipped; the sigma terms remain."""
    p = _variety(fko, 17, 24, 0, 80, 16, 6, 5)
    p.ekq = p.ekq + 100.0
    ref = fko.ff_full(p)
    assert np.max(np.abs(ref["asxDtemp"])) == 0
    for memory in ("device", "host"):
        got = _np(fk.ff_full_problem(_to_dev(p) if memory == "device" else p))
        for k in ("achDtemp", "achDtemp_cor", "achDtemp_corb", "ach2Dtemp", "schDt"):
            assert _rel(got[k], ref[k]) <= RTOL, (memory, k)
        assert np.max(np.abs(got["asxDtemp"])) == 0
        assert abs(got["achsDtemp"] - ref["achsDtemp"]) <= RTOL * abs(ref["achsDtemp"])


def test_skipped_columns_and_achs_freq(fk, fko):
    p = _variety(fko, 11, 20, 6, 64, 16, 6, 8)
    inv = p.inv_igp_index.copy()
    inv[[2, 5]] = inv[[5, 2]]
    inv[7] = p.ncouls                     # indigp == ncouls -> igp = ncouls - 1
    ind = p.indinv.copy()
    ind[inv[9]] = -1                      # skipped column
    p.inv_igp_index, p.indinv = inv, ind
    p.achs_freq = 3
    ref = fko.ff_full(p)
    got = _np(fk.ff_full_problem(_to_dev(p)))
    _check(got, ref, RTOL, "skip map")


def test_host_memory_same_bits_and_workspace(fk, fko):
    torch = _torch()
    p = _variety(fko, 21, 40, 7, 120, 24, 7, 9, random_axes=True)
    a = _np(fk.ff_full_problem(_to_dev(p)))
    b = _np(fk.ff_full_problem(_to_dev(p)))
    h = fk.ff_full_problem(p)             # host arrays: uploads, I_epsA only over the frequencies read
    for k in OUTS + ("schDt",):
        assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(a[k], h[k]), k
    n = fk.ff_full_workspace_size(p.nb, p.nv, p.ncouls, p.ngpown, p.nfreq, p.nfreqeval, p.ekq, p.dFreqGrid,
                                  freqevalmin=p.freqevalmin, freqevalstep=p.freqevalstep)
    ws = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        c = fk.ff_full_problem(_to_dev(p), stream=s, workspace=ws, time_kernel=True, want_schDt=False)
    s.synchronize()
    c = _np(c)
    assert c["kernel_ms"] > 0 and c["schDt"] is None
    for k in OUTS:
        assert np.array_equal(a[k], c[k]), k


def test_full_size_program_inputs_scale(fk, fko):
    """256/8/4096/4096/32/10 with the program's constant inputs
    (ffkernel.deslippe.best.f90:111-206; aqs = 0.5+0.5i, I_epsR = 0.5+0.5i,
    I_epsA = 0.5-0.5i, vcoul = 1): every output is G*C times that of any
    other (G, C), so the oracle's 8 x 8 result scaled by 4096*4096/64 is the
    expected value."""
    torch = _torch()
    NB, NV, C, G, NF, NE = 256, 8, 4096, 4096, 32, 10
    ekq, grid, pref, brd = fko.ff_program_axes(NB, NF)

    def const_problem(c, g):
        z = np.full((NB, c), 0.5 + 0.5j)
        inv = np.arange(g, dtype=np.int32)
        ind = np.arange(c + 1, dtype=np.int32)
        ind[c] = c - 1
        return fko.FfFullProblem(NB, NV, c, g, NF, NE, z, z.copy(), np.full((NF, g, c), 0.5 + 0.5j),
                                 np.full((NF, g, c), 0.5 - 0.5j), np.ones(c), inv, ind, ekq, grid, pref, brd)

    small = fko.ff_full(const_problem(8, 8))
    scale = (G * C) / 64.0
    d = torch.device("cuda:0")
    aq = torch.full((NB, C), 0.5 + 0.5j, dtype=torch.complex128, device=d)
    IR = torch.full((NF, G, C), 0.5 + 0.5j, dtype=torch.complex128, device=d)
    IA = torch.full((NF, G, C), 0.5 - 0.5j, dtype=torch.complex128, device=d)
    v = torch.ones(C, dtype=torch.float64, device=d)
    inv = np.arange(G, dtype=np.int32)
    ind = np.arange(C + 1, dtype=np.int32)
    ind[C] = C - 1
    print("ff_full full-size constant inputs resident", flush=True)
    for prec in ("fp64", "mixed"):
        got = _np(fk.ff_kernel_full(NB, NV, C, G, NF, NE, aq, aq, IR, IA, v, inv, ind, ekq, grid, pref, brd,
                                    time_kernel=True, precision=prec))
        ref = {k: small[k] * scale for k in OUTS + ("schDt",)}
        ref["achsDtemp"] = small["achsDtemp"] * scale
        print(f"  {prec}: {got['kernel_ms']:.3f} ms", flush=True)
        # 0.5 +- 0.5i is exact in BF16, so the split path is exact here too
        _check(got, ref, RTOL, f"full size {prec}")
    del IR, IA
    torch.cuda.empty_cache()


def test_full_size_random_mixed_vs_fp64(fk, fko):
    """256/8/4096/4096/32/10, random inputs generated on the device: the
    split-BF16 result against the FP64 result (both from the device; FP64 is
    held to 1e-12 against the oracle above)."""
    torch = _torch()
    NB, NV, C, G, NF, NE = 256, 8, 4096, 4096, 32, 10
    ekq, grid, pref, brd = fko.ff_program_axes(NB, NF)
    d = torch.device("cuda:0")
    aqsm = torch.empty(NB, C, dtype=torch.complex128, device=d)
    aqsn = torch.empty(NB, C, dtype=torch.complex128, device=d)
    IR = torch.empty(NF, G, C, dtype=torch.complex128, device=d)
    IA = torch.empty(NF, G, C, dtype=torch.complex128, device=d)
    v = torch.empty(C, dtype=torch.float64, device=d)
    fk.fill_complex(aqsm, 4, 1, (-0.5, 0.5), (-0.5, 0.5))
    fk.fill_complex(aqsn, 4, 2, (-0.5, 0.5), (-0.5, 0.5))
    fk.fill_complex(IR, 4, 6, (-0.5, 0.5), (-0.5, 0.5))
    fk.fill_complex(IA, 4, 7, (-0.5, 0.5), (-0.5, 0.5))
    fk.fill_real(v, 4, 5, 0.1, 10.0)
    inv = np.arange(G, dtype=np.int32)
    ind = np.arange(C + 1, dtype=np.int32)
    ind[C] = C - 1
    print("ff_full full-size random inputs filled", flush=True)
    res = {}
    for prec in ("fp64", "mixed"):
        res[prec] = _np(fk.ff_kernel_full(NB, NV, C, G, NF, NE, aqsm, aqsn, IR, IA, v, inv, ind, ekq, grid, pref,
                                          brd, time_kernel=True, precision=prec))
        print(f"  {prec}: {res[prec]['kernel_ms']:.3f} ms", flush=True)
    _check(res["mixed"], res["fp64"], MIXED_RTOL, "full size mixed vs fp64")
    del IR, IA
    torch.cuda.empty_cache()


# ---------------------------------------------------------------------------
# semantics FK_FF_CPP_PORT / FK_FF_CPP_PORT_ZERO_SCHDT: the C++ port
# FFKer_cori/ffKernel.cpp as its main calls it
# ---------------------------------------------------------------------------
def _fixture(fko, name):
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", f"ff_full_{name}.npz"))
    nb, nv, C, G, nF, ne = (int(x) for x in d["sizes"])
    pz, fmin, fstep, occ = (float(x) for x in d["scalars"])
    p = fko.FfFullProblem(nb, nv, C, G, nF, ne, d["aqsm"], d["aqsn"], d["I_epsR"], d["I_epsA"], d["vcoul"],
                          d["inv_igp_index"], d["indinv"], d["ekq"], d["dFreqGrid"], d["pref"], d["dFreqBrd"],
                          pz, fmin, fstep, occ, achs_freq=1)
    return p, d


@pytest.mark.parametrize("name", ["program_axes", "negative_axes"])
def test_cpp_port_matches_reference_functions(fk, fko, name):
    """ff_full_<name>.npz holds what the C++ port's own kernel functions
    (compiled from FFKer_cori/ffKernel.cpp, one thread) return for these
    inputs with schDt_matrix = the contraction of I_epsR, which is what the
    device evaluates with under FK_FF_CPP_PORT: the program's axes, and a
    non-uniform grid with freqevalmin -1.75, freqevalstep 2.5 (which the port
    truncates to -1 and 2)."""
    p, d = _fixture(fko, name)
    print(f"ff_full cpp port vs reference functions, {name}", flush=True)
    got = _np(fk.ff_full_problem(_to_dev(p), semantics=fk.FK_FF_CPP_PORT))
    for k in OUTS:
        ref = d[k]
        if np.max(np.abs(ref)) == 0:
            assert np.max(np.abs(got[k])) == 0, k
        else:
            assert _rel(got[k], ref) <= RTOL, (k, _rel(got[k], ref))
    assert abs(got["achsDtemp"] - complex(d["achsDtemp"])) <= RTOL * abs(complex(d["achsDtemp"]))


@pytest.mark.parametrize("case", [("seeded", 0), ("seeded", 1), ("variety", 3), ("variety", 4), ("variety", 5),
                                  ("variety", 6)])
def test_cpp_port_vs_oracle(fk, fko, case):
    """The port's arithmetic on inputs that exercise it: a non-uniform grid,
    band energies that put wx' on both sides of zero, freqevalmin and
    freqevalstep that the port truncates to int (0.25, 1.5), every band
    occupied with two grid frequencies, no band occupied with one evaluation
    frequency.  Against the
    oracle's variant 1 (pinned to the compiled port) with the oracle's own
    contraction as schDt_matrix (FK_FF_CPP_PORT) and with zeros
    (FK_FF_CPP_PORT_ZERO_SCHDT, the C++ program as it runs)."""
    kind, seed = case
    if kind == "seeded":
        p, _ = _fixture(fko, "seeded")
        if seed:
            p.freqevalmin, p.freqevalstep = -1.75, 2.5
    else:
        p = _variety(fko, seed, *CASES[seed - 1][1:7], fmin=0.6, step=1.3, random_axes=True)
        p.achs_freq = 1
    print(f"ff_full cpp port case {case}", flush=True)
    S = fko.ff_full(p)["schDt"]
    for sem, schDt_in in ((fk.FK_FF_CPP_PORT, S), (fk.FK_FF_CPP_PORT_ZERO_SCHDT, None)):
        ref = fko.ff_full(p, variant=1, schDt_in=schDt_in)
        for memory in ("device", "host"):
            got = _np(fk.ff_full_problem(_to_dev(p) if memory == "device" else p, semantics=sem))
            for k in OUTS:
                if np.max(np.abs(ref[k])) == 0:
                    assert np.max(np.abs(got[k])) == 0, (sem, memory, k)
                else:
                    assert _rel(got[k], ref[k]) <= RTOL, (sem, memory, k, _rel(got[k], ref[k]))
            assert _rel(got["schDt"], S) <= RTOL
            assert abs(got["achsDtemp"] - ref["achsDtemp"]) <= RTOL * abs(ref["achsDtemp"])


def test_cpp_port_mixed(fk, fko):
    p, _ = _fixture(fko, "seeded")
    print("ff_full cpp port mixed", flush=True)
    ref = fko.ff_full(p, variant=1, schDt_in=fko.ff_full(p)["schDt"])
    got = _np(fk.ff_full_problem(_to_dev(p), semantics=fk.FK_FF_CPP_PORT, precision="mixed"))
    errs = {k: _rel(got[k], ref[k]) for k in OUTS if np.max(np.abs(ref[k])) > 0}
    print("  " + " ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    assert max(errs.values()) <= MIXED_RTOL
