"""GPU accuracy of the mixed-precision GPP path (complex64 inputs, band
contraction on int8 MFMA by an exact-accumulation split, FP64 elsewhere)
against the FP64 oracle evaluated on the same complex64 values.  The bar is
the 1e-8 relative error the mixed form is specified to; the FP32 MFMA
variant (config 2) is held to 1e-5."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIXED_RTOL = 1e-8


def _torch():
    import torch
    assert torch.cuda.is_available(), "no ROCm device visible"
    return torch


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _round64(p):
    for k in ("aqsm", "aqsn", "I_eps", "wtilde"):
        setattr(p, k, getattr(p, k).astype(np.complex64).astype(np.complex128))
    return p


def _run(fk, p, memory="device", **kw):
    torch = _torch()
    if memory == "host":
        h = {k: np.ascontiguousarray(getattr(p, k).astype(np.complex64)) for k in ("aqsm", "aqsn", "I_eps", "wtilde")}
        ach, asx, _ = fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, h["aqsm"], h["aqsn"], h["I_eps"],
                                    h["wtilde"], p.vcoul, p.inv_igp_index, p.indinv, precision="mixed", **kw)
        return ach, asx
    f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0").to(dt)  # noqa: E731
    t = {k: f(getattr(p, k), torch.complex64) for k in ("aqsm", "aqsn", "I_eps", "wtilde")}
    v = f(p.vcoul, torch.float64)
    ach, asx, _ = fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, t["aqsm"], t["aqsn"], t["I_eps"],
                                t["wtilde"], v, p.inv_igp_index, p.indinv, precision="mixed", **kw)
    torch.cuda.synchronize()
    return ach.cpu().numpy(), asx.cpu().numpy()


@pytest.mark.parametrize("shape", [(128, 2, 8192, 409), (7, 3, 100, 9), (1, 1, 1, 1), (33, 0, 257, 17),
                                   (20, 20, 640, 31), (300, 16, 1000, 70)])
def test_mixed_vs_fp64_oracle(fk, fko, shape):
    nb, nv, C, G = shape
    print(f"mixed shape {shape}", flush=True)
    p = _round64(fko.gpp_random(nb, nv, C, G, seed=5, wtilde_re=(3.0, 5.0)))
    ra, rs, _ = fko.gpp(p)
    ach, asx = _run(fk, p)
    ea, es = _rel(ach, ra), _rel(asx, rs) if nv else 0.0
    print(f"  rel err achtemp {ea:.2e} asxtemp {es:.2e}", flush=True)
    assert ea <= MIXED_RTOL and es <= MIXED_RTOL
    if nv == 0:
        assert np.all(asx == 0)


@pytest.mark.parametrize("nw", [1, 5, 16])
def test_mixed_frequency_counts(fk, fko, nw):
    p = _round64(fko.gpp_random(40, 3, 700, 35, nw=nw, seed=nw))
    ra, rs, _ = fko.gpp(p)
    ach, asx = _run(fk, p)
    assert _rel(ach, ra) <= MIXED_RTOL and _rel(asx, rs) <= MIXED_RTOL


@pytest.mark.parametrize("consts", [(1e-4, 0.5, 1e3, 0.5), (1e-8, 0.04, 1e7, 2.0)])
def test_mixed_thresholds_and_evaluation_axis(fk, fko, consts):
    """The int8-split path with other thresholds and an evaluation axis with
    a negative frequency and points inside the wtilde range."""
    p = _round64(fko.gpp_random(48, 5, 640, 37, nw=4, seed=77, wtilde_re=(0.5, 5.0)))
    p.to1, p.limittwo, p.limitone, p.occ = consts
    p.wx = np.array([-1.3, 0.7, 2.9, 4.4])
    print(f"mixed thresholds {consts}", flush=True)
    ra, rs, _ = fko.gpp(p)
    ach, asx = _run(fk, p, to1=p.to1, limittwo=p.limittwo, limitone=p.limitone, occ=p.occ)
    ea, es = _rel(ach, ra), _rel(asx, rs)
    print(f"  errors {ea:.2e} {es:.2e}", flush=True)
    assert ea <= MIXED_RTOL and es <= MIXED_RTOL


def test_mixed_host_memory_mask_off_and_skipped_columns(fk, fko):
    p = _round64(fko.gpp_random(24, 2, 512, 40, seed=8, wtilde_re=(3.0, 5.0)))
    ind = p.indinv.copy()
    ind[p.inv_igp_index[3]] = -1
    p.indinv = ind
    for mask in (True, False):
        ra, rs, _ = fko.gpp(p, mask_on=mask)
        ach, asx = _run(fk, p, memory="host", mask=mask)
        assert _rel(ach, ra) <= MIXED_RTOL and _rel(asx, rs) <= MIXED_RTOL


def test_mixed_wide_dynamic_range(fk, fko):
    """Columns of aqsn and rows of the coefficients spanning 2^-40 .. 2^10:
    every row and column carries its own power-of-two scale."""
    p = fko.gpp_random(64, 4, 1024, 51, seed=12)
    rng = np.random.default_rng(3)
    p.aqsn = p.aqsn * np.exp2(rng.integers(-40, 10, size=(1, p.ncouls)))
    p.aqsm = p.aqsm * np.exp2(rng.integers(-30, 5, size=(1, p.ncouls)))
    p = _round64(p)
    ra, rs, _ = fko.gpp(p)
    ach, asx = _run(fk, p)
    assert _rel(ach, ra) <= MIXED_RTOL and _rel(asx, rs) <= MIXED_RTOL


def test_mixed_fp32_mfma_variant(fk, fko):
    p = _round64(fko.gpp_random(64, 4, 4096, 205, seed=2, wtilde_re=(3.0, 5.0)))
    ra, rs, _ = fko.gpp(p)
    ach, asx = _run(fk, p, config=2)
    assert _rel(ach, ra) <= 1e-5 and _rel(asx, rs) <= 1e-5


def test_mixed_full_size_vs_fp64_kernel(fk, fko):
    """512/2/32768/20 with wtilde re ~ U(3,5): the mixed result against the
    FP64 kernel (itself held to 1e-12 of the oracle) on the same complex64
    values, and bitwise determinism."""
    torch = _torch()
    nb, nv, C, G = 512, 2, 32768, 1638
    t64, t32 = {}, {}
    for name, rows, re in (("aqsm", nb, (-0.5, 0.5)), ("aqsn", nb, (-0.5, 0.5)), ("I_eps", G, (-0.5, 0.5)),
                           ("wtilde", G, (3.0, 5.0))):
        t32[name] = torch.empty(rows, C, dtype=torch.complex64, device="cuda:0")
        im = (-0.01, 0.01) if name == "wtilde" else (-0.5, 0.5)
        fk.fill_complex(t32[name], 1, fko.STREAM[name], re, im)
        t64[name] = t32[name].to(torch.complex128)
    v = torch.empty(C, dtype=torch.float64, device="cuda:0")
    fk.fill_real(v, 1, fko.STREAM["vcoul"], 0.1, 10.0)
    inv, indinv = fko.index_maps(C, G, "fortran")
    wx = fko.wx_array(3)
    args = (inv, indinv)
    a64, s64, _ = fk.gpp_kernel(nb, nv, C, G, wx, t64["aqsm"], t64["aqsn"], t64["I_eps"], t64["wtilde"], v, *args)
    a32, s32, _ = fk.gpp_kernel(nb, nv, C, G, wx, t32["aqsm"], t32["aqsn"], t32["I_eps"], t32["wtilde"], v, *args,
                                precision="mixed")
    b32, c32, _ = fk.gpp_kernel(nb, nv, C, G, wx, t32["aqsm"], t32["aqsn"], t32["I_eps"], t32["wtilde"], v, *args,
                                precision="mixed")
    torch.cuda.synchronize()
    ea, es = _rel(a32.cpu().numpy(), a64.cpu().numpy()), _rel(s32.cpu().numpy(), s64.cpu().numpy())
    print(f"full-size mixed rel err: achtemp {ea:.2e} asxtemp {es:.2e}", flush=True)
    assert ea <= MIXED_RTOL and es <= MIXED_RTOL
    assert torch.equal(a32, b32) and torch.equal(s32, c32)
