"""GPU parity of gpp_kernel (through the C ABI) against the CPU oracle, the
reference's golden vectors and, at the full 512/2/32768/20 size, against
size-independent properties.  Bar: relative error <= 1e-12 for FP64."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _torch():
    import torch
    assert torch.cuda.is_available(), "no ROCm device visible"
    return torch


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _assert_close(got, ref, rtol=RTOL, what=""):
    got, ref = np.asarray(got), np.asarray(ref)
    scale = np.max(np.abs(ref))
    for g, r in zip(got, ref):
        assert abs(g - r) <= rtol * abs(r) + 1e-3 * rtol * scale, f"{what}: got {g} want {r}"


def _dev(p, cdt=None):
    torch = _torch()
    cdt = cdt or torch.complex128
    d = torch.device("cuda:0")
    f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(d).to(dt)  # noqa: E731
    return dict(aqsm=f(p.aqsm, cdt), aqsn=f(p.aqsn, cdt), I_eps=f(p.I_eps, cdt), wtilde=f(p.wtilde, cdt),
                vcoul=f(p.vcoul, torch.float64))


def _run_dev(fk, p, **kw):
    t = _dev(p)
    ach, asx, ms = fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, t["aqsm"], t["aqsn"], t["I_eps"],
                                 t["wtilde"], t["vcoul"], p.inv_igp_index, p.indinv, to1=p.to1,
                                 limittwo=p.limittwo, limitone=p.limitone, occ=p.occ, **kw)
    _torch().cuda.synchronize()
    return ach.cpu().numpy(), asx.cpu().numpy(), ms


def _from_npz(fko, d):
    return fko.GppProblem(int(d["nb"]), int(d["nv"]), int(d["ncouls"]), int(d["ngpown"]), d["wx"], d["aqsm"],
                          d["aqsn"], d["I_eps"], d["wtilde"], d["vcoul"], d["inv_igp_index"], d["indinv"],
                          occ=float(d["occ"]))


# accumulator register r of lane l holds row acc_row(l, r), column l % 16
ACC_ROW = {0: lambda l, r: 4 * r + l // 16,       # FP64 16x16x4
           1: lambda l, r: 4 * (l // 16) + r}     # FP32 16x16x4


@pytest.mark.parametrize("kind", [0, 1])
def test_mfma_result_layout(fk, kind):
    out = fk.mfma_probe(kind)
    A = np.array([[1.0 + i + 16.0 * k for k in range(4)] for i in range(16)])
    B = np.array([[1.0 + 3.0 * j - k for j in range(16)] for k in range(4)])
    Cm = A @ B
    bad = []
    for lane in range(64):
        for r in range(4):
            if out[lane, r] != Cm[ACC_ROW[kind](lane, r), lane % 16]:
                cands = [(int(i), int(j)) for i, j in zip(*np.nonzero(Cm == out[lane, r]))]
                bad.append((lane, r, cands))
    assert not bad, f"kind {kind}: {len(bad)} mismatches, first: {bad[:6]}"


def test_device_fill_matches_host_generator(fk, fko):
    torch = _torch()
    z = torch.empty(1000, dtype=torch.complex128, device="cuda:0")
    fk.fill_complex(z, 5, fko.STREAM["wtilde"], (0.5, 2.0), (-0.01, 0.01))
    r = torch.empty(777, dtype=torch.float64, device="cuda:0")
    fk.fill_real(r, 5, fko.STREAM["vcoul"], 0.1, 10.0)
    z32 = torch.empty(1000, dtype=torch.complex64, device="cuda:0")
    fk.fill_complex(z32, 5, fko.STREAM["wtilde"], (0.5, 2.0), (-0.01, 0.01))
    torch.cuda.synchronize()
    zh = fko.fill_complex(1000, 5, fko.STREAM["wtilde"], (0.5, 2.0), (-0.01, 0.01))
    assert np.array_equal(z.cpu().numpy(), zh)
    assert np.array_equal(r.cpu().numpy(), fko.fill_real(777, 5, fko.STREAM["vcoul"], 0.1, 10.0))
    assert np.array_equal(z32.cpu().numpy(), zh.astype(np.complex64))


def test_complex_helpers_device_match_host(fk):
    rng = np.random.default_rng(1)
    a = rng.normal(size=300) + 1j * rng.normal(size=300)
    b = rng.normal(size=300) + 1j * rng.normal(size=300)
    s = rng.uniform(0.5, 2.0, 300)
    acc = rng.normal(size=300) + 1j * rng.normal(size=300)
    for op in fk.COP:
        h = fk.complex_apply(op, a=a, b=b, s=s, out=acc, on_device=False)
        d = fk.complex_apply(op, a=a, b=b, s=s, out=acc, on_device=True)
        assert np.allclose(d, h, rtol=4e-16, atol=1e-15), op


@pytest.mark.parametrize("case", ["random", "nearpole", "branch3", "skipmap"])
@pytest.mark.parametrize("memory", ["host", "device"])
def test_gpp_golden_vectors(fk, fko, case, memory):
    d = np.load(os.path.join(GOLDEN, f"gpp_{case}.npz"))
    p = _from_npz(fko, d)
    if memory == "host":
        ach, asx, _ = fk.gpp_problem(p)
    else:
        ach, asx, _ = _run_dev(fk, p)
    _assert_close(ach, d["achtemp"], what=f"{case} achtemp")
    _assert_close(asx, d["asxtemp"], what=f"{case} asxtemp")


def test_gpp_reference_program_outputs(fk, fko):
    d = np.load(os.path.join(GOLDEN, "gpp_main.npz"))
    for (nb, nv, C, npg), ref in zip(d["shapes"], d["achtemp"]):
        p = fko.gpp_constant(int(nb), int(nv), int(C), int(npg), map_kind="cpp")
        ach, asx, _ = fk.gpp_problem(p)
        _assert_close(ach, ref, rtol=1e-13, what=f"main {nb},{nv},{C},{npg}")
        _, ref_asx = fko.gpp_closed_form(p.nb, p.nv, p.ncouls, p.ngpown, p.wx)
        _assert_close(asx, ref_asx, rtol=1e-13, what="asx closed form")


@pytest.mark.parametrize("shape", [
    (128, 2, 8192, 409),      # the 128/2/8192/20 workload
    (7, 3, 100, 9),           # ragged: no dimension a tile multiple
    (1, 1, 1, 1),
    (33, 0, 257, 17),         # no occupied band
    (20, 20, 640, 31),        # every band occupied
])
def test_gpp_vs_oracle(fk, fko, shape):
    nb, nv, C, G = shape
    p = fko.gpp_random(nb, nv, C, G, seed=0)
    ref_ach, ref_asx, _ = fko.gpp(p, nthreads=0)
    ach, asx, _ = _run_dev(fk, p)
    _assert_close(ach, ref_ach, what=f"{shape} achtemp")
    _assert_close(asx, ref_asx, what=f"{shape} asxtemp")


@pytest.mark.parametrize("nw", [1, 5, 11, 16])
def test_gpp_frequency_counts(fk, fko, nw):
    p = fko.gpp_random(9, 3, 300, 12, nw=nw, seed=nw)
    ref_ach, ref_asx, _ = fko.gpp(p)
    ach, asx, _ = _run_dev(fk, p)
    _assert_close(ach, ref_ach, what=f"nw={nw}")
    _assert_close(asx, ref_asx, what=f"nw={nw}")


@pytest.mark.parametrize("mask", [True, False])
@pytest.mark.parametrize("consts", [(1e-4, 0.5, 1e3, 0.5), (1e-8, 0.04, 1e7, 2.0)])
def test_gpp_thresholds_and_evaluation_axis(fk, fko, consts, mask):
    """Other to1 / limittwo / limitone / occ than the program's, and an
    evaluation axis with a negative frequency and points inside the wtilde
    range (near-pole terms in every branch), with the canonical mask and
    without it (the GPU variant's achtemp)."""
    p = fko.gpp_random(48, 5, 640, 37, nw=4, seed=77, wtilde_re=(0.5, 5.0))
    p.to1, p.limittwo, p.limitone, p.occ = consts
    p.wx = np.array([-1.3, 0.7, 2.9, 4.4])
    print(f"gpp thresholds {consts} mask {mask}", flush=True)
    ref_ach, ref_asx, _ = fko.gpp(p, mask_on=mask)
    ach, asx, _ = _run_dev(fk, p, mask=mask)
    _assert_close(ach, ref_ach, what=f"achtemp {consts}")
    _assert_close(asx, ref_asx, what=f"asxtemp {consts}")


def test_gpp_mask_off_semantics(fk, fko):
    d = np.load(os.path.join(GOLDEN, "gpp_nearpole.npz"))
    p = _from_npz(fko, d)
    ref_ach, ref_asx, _ = fko.gpp(p, mask_on=False)
    ach, asx, _ = _run_dev(fk, p, mask=False)
    _assert_close(ach, ref_ach, what="mask off")
    _assert_close(asx, ref_asx, what="mask off asx")


def test_gpp_tile_configs_agree(fk, fko):
    p = fko.gpp_random(40, 4, 2000, 100, seed=8)
    ref_ach, ref_asx, _ = fko.gpp(p)
    for cfg in (1, 2):
        ach, asx, _ = _run_dev(fk, p, config=cfg)
        _assert_close(ach, ref_ach, what=f"config {cfg}")
        _assert_close(asx, ref_asx, what=f"config {cfg}")


def test_gpp_workspace_stream_and_timer(fk, fko):
    torch = _torch()
    p = fko.gpp_random(16, 2, 1024, 51, seed=9)
    ref_ach, _, _ = fko.gpp(p)
    t = _dev(p)
    nbytes = fk.gpp_workspace_size(p.nb, p.nv, p.ncouls, p.ngpown, p.nw)
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ach, asx, ms = fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, t["aqsm"], t["aqsn"], t["I_eps"],
                                     t["wtilde"], t["vcoul"], p.inv_igp_index, p.indinv, stream=s,
                                     workspace=ws, time_kernel=True)
    s.synchronize()
    assert ms is not None and ms > 0
    _assert_close(ach.cpu().numpy(), ref_ach, what="stream+workspace")
    small = torch.empty(max(nbytes // 2, 1), dtype=torch.uint8, device="cuda:0")
    with pytest.raises(fk.FkError) as e:
        fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, t["aqsm"], t["aqsn"], t["I_eps"], t["wtilde"],
                      t["vcoul"], p.inv_igp_index, p.indinv, workspace=small)
    assert e.value.code == fk.FK_ERR_WORKSPACE


# ---------------------------------------------------------------------------
# full size 512/2/32768/20 (ngpown 1638): closed form + properties
# ---------------------------------------------------------------------------
FULL = (512, 2, 32768, 1638)


def test_full_size_closed_form(fk, fko):
    torch = _torch()
    nb, nv, C, G = FULL
    dev = "cuda:0"
    z = torch.full((nb, C), 0.5 + 0.5j, dtype=torch.complex128, device=dev)
    w = torch.full((G, C), 0.5 + 0.5j, dtype=torch.complex128, device=dev)
    v = torch.ones(C, dtype=torch.float64, device=dev)
    inv, indinv = fko.index_maps(C, G, "fortran")
    wx = fko.wx_array(3)
    ach, asx, _ = fk.gpp_kernel(nb, nv, C, G, wx, z, z, w, w, v, inv, indinv)
    torch.cuda.synchronize()
    ref_ach, ref_asx = fko.gpp_closed_form(nb, nv, C, G, wx)
    _assert_close(ach.cpu().numpy(), ref_ach, rtol=1e-13, what="full closed form")
    _assert_close(asx.cpu().numpy(), ref_asx, rtol=1e-13, what="full closed form asx")


@pytest.fixture(scope="module")
def full_random(fk, fko):
    torch = _torch()
    nb, nv, C, G = FULL
    dev = "cuda:0"
    t = {}
    for name, n, re in (("aqsm", nb * C, (-0.5, 0.5)), ("aqsn", nb * C, (-0.5, 0.5)),
                        ("I_eps", G * C, (-0.5, 0.5)), ("wtilde", G * C, (0.5, 2.0))):
        t[name] = torch.empty(n, dtype=torch.complex128, device=dev)
        im = (-0.01, 0.01) if name == "wtilde" else (-0.5, 0.5)
        fk.fill_complex(t[name], 1, fko.STREAM[name], re, im)
    t["vcoul"] = torch.empty(C, dtype=torch.float64, device=dev)
    fk.fill_real(t["vcoul"], 1, fko.STREAM["vcoul"], 0.1, 10.0)
    for k in ("aqsm", "aqsn"):
        t[k] = t[k].view(nb, C)
    for k in ("I_eps", "wtilde"):
        t[k] = t[k].view(G, C)
    torch.cuda.synchronize()
    inv, indinv = fko.index_maps(C, G, "fortran")
    return t, inv, indinv, fko.wx_array(3)


def _full_call(fk, t, inv, indinv, wx, nb=None, nv=None, lo=0, **kw):
    nb_all, nv_all, C, G = FULL
    nb = nb_all if nb is None else nb
    nv = nv_all if nv is None else nv
    ach, asx, ms = fk.gpp_kernel(nb, nv, C, G, wx, t["aqsm"][lo:lo + nb], t["aqsn"][lo:lo + nb], t["I_eps"],
                                 t["wtilde"], t["vcoul"], inv, indinv, **kw)
    _torch().cuda.synchronize()
    return ach.cpu().numpy(), asx.cpu().numpy()


def test_full_size_deterministic_and_linear(fk, full_random):
    t, inv, indinv, wx = full_random
    a1, s1 = _full_call(fk, t, inv, indinv, wx)
    a2, s2 = _full_call(fk, t, inv, indinv, wx)
    assert np.array_equal(a1, a2) and np.array_equal(s1, s2)
    t2 = dict(t, vcoul=t["vcoul"] * 2.0)         # power-of-two scaling is exact
    a3, s3 = _full_call(fk, t2, inv, indinv, wx)
    assert np.array_equal(a3, 2 * a1) and np.array_equal(s3, 2 * s1)


def test_full_size_band_additivity(fk, full_random):
    """achtemp over 512 bands = bands [0,256) (2 occupied) + bands [256,512) (none)."""
    t, inv, indinv, wx = full_random
    a, s = _full_call(fk, t, inv, indinv, wx)
    a_lo, s_lo = _full_call(fk, t, inv, indinv, wx, nb=256, nv=2, lo=0)
    a_hi, s_hi = _full_call(fk, t, inv, indinv, wx, nb=256, nv=0, lo=256)
    _assert_close(a_lo + a_hi, a, rtol=1e-13, what="band additivity")
    _assert_close(s_lo, s, rtol=1e-14, what="asx from occupied bands only")
    assert np.all(s_hi == 0)


def test_full_size_column_additivity(fk, full_random):
    """Skipping a set of columns (negative indinv) splits achtemp additively."""
    t, inv, indinv, wx = full_random
    C = FULL[2]
    a, s = _full_call(fk, t, inv, indinv, wx)
    even = indinv.copy()
    odd = indinv.copy()
    even[inv[1::2]] = -1      # keep even my_igp
    odd[inv[0::2]] = -1       # keep odd my_igp
    assert C not in inv
    a_e, s_e = _full_call(fk, t, inv, even, wx)
    a_o, s_o = _full_call(fk, t, inv, odd, wx)
    _assert_close(a_e + a_o, a, rtol=1e-13, what="column additivity")
    _assert_close(s_e + s_o, s, rtol=1e-13, what="column additivity asx")


def test_full_shape_vs_oracle_two_bands(fk, fko, full_random):
    """The oracle at the full ngpown x ncouls shape for the two occupied bands."""
    t, inv, indinv, wx = full_random
    nb_all, nv, C, G = FULL
    p = fko.GppProblem(2, 2, C, G, wx, t["aqsm"][:2].cpu().numpy(), t["aqsn"][:2].cpu().numpy(),
                       t["I_eps"].cpu().numpy(), t["wtilde"].cpu().numpy(), t["vcoul"].cpu().numpy(), inv, indinv)
    ref_a, ref_s, _ = fko.gpp(p, nthreads=0)
    a, s = _full_call(fk, t, inv, indinv, wx, nb=2, nv=2)
    _assert_close(a, ref_a, what="full shape 2 bands")
    _assert_close(s, ref_s, what="full shape 2 bands asx")


# ---------------------------------------------------------------------------
# mixed precision (complex64 storage)
# ---------------------------------------------------------------------------
def test_gpp_mixed_vs_fp64_oracle(fk, fko):
    """Mixed precision on complex64 inputs vs the FP64 oracle on the same
    (complex64-rounded) values; the bound documents the measured error."""
    torch = _torch()
    p = fko.gpp_random(64, 4, 4096, 205, seed=2, wtilde_re=(3.0, 5.0))
    for k in ("aqsm", "aqsn", "I_eps", "wtilde"):
        setattr(p, k, getattr(p, k).astype(np.complex64).astype(np.complex128))
    ref_a, ref_s, _ = fko.gpp(p, nthreads=0)
    t = _dev(p, torch.complex64)
    ach, asx, _ = fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, t["aqsm"], t["aqsn"], t["I_eps"],
                                t["wtilde"], t["vcoul"], p.inv_igp_index, p.indinv, precision="mixed")
    torch.cuda.synchronize()
    ea, es = _rel(ach.cpu().numpy(), ref_a), _rel(asx.cpu().numpy(), ref_s)
    print(f"mixed-precision relative error vs FP64 oracle: achtemp {ea:.3e} asxtemp {es:.3e}", flush=True)
    assert ea < 1e-5 and es < 1e-5
