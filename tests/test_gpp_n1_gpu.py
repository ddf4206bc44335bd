"""GPU parity of gpp_kernel's per-band output acht_n1_loc (the per-band
achtemp at frequency index 2, BGW-Kernels/gppKer.cpp:424) against the
reference's own values in the golden vectors, the CPU oracle, and the
invariant sum_n1 acht_n1_loc[n1] = achtemp[2].  Bar: relative error <= 1e-12."""
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


def _close(got, ref, rtol=RTOL, what=""):
    got, ref = np.asarray(got), np.asarray(ref)
    scale = np.max(np.abs(ref))
    bad = np.abs(got - ref) > rtol * np.abs(ref) + 1e-3 * rtol * scale
    assert not bad.any(), f"{what}: {int(bad.sum())} of {bad.size} off, first at {int(np.argmax(bad))}"


def _dev(p, cdt=None):
    torch = _torch()
    cdt = cdt or torch.complex128
    f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0").to(dt)  # noqa: E731
    return dict(aqsm=f(p.aqsm, cdt), aqsn=f(p.aqsn, cdt), I_eps=f(p.I_eps, cdt), wtilde=f(p.wtilde, cdt),
                vcoul=f(p.vcoul, torch.float64))


def _run(fk, p, memory="device", **kw):
    if memory == "host":
        ach, asx, _, n1 = fk.gpp_problem(p, acht_n1=True, **kw)
        return ach, asx, n1
    t = _dev(p)
    ach, asx, _, n1 = fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, t["aqsm"], t["aqsn"], t["I_eps"],
                                    t["wtilde"], t["vcoul"], p.inv_igp_index, p.indinv, to1=p.to1,
                                    limittwo=p.limittwo, limitone=p.limitone, occ=p.occ, acht_n1=True, **kw)
    _torch().cuda.synchronize()
    return ach.cpu().numpy(), asx.cpu().numpy(), n1.cpu().numpy()


@pytest.mark.parametrize("case", ["random", "nearpole", "branch3", "skipmap"])
@pytest.mark.parametrize("memory", ["host", "device"])
def test_n1_golden_vectors(fk, fko, case, memory):
    d = np.load(os.path.join(GOLDEN, f"gpp_{case}.npz"))
    p = fko.GppProblem(int(d["nb"]), int(d["nv"]), int(d["ncouls"]), int(d["ngpown"]), d["wx"], d["aqsm"],
                       d["aqsn"], d["I_eps"], d["wtilde"], d["vcoul"], d["inv_igp_index"], d["indinv"],
                       occ=float(d["occ"]))
    ach, asx, n1 = _run(fk, p, memory)
    _close(n1, d["acht_n1_loc"], what=f"{case} acht_n1_loc")
    _close(ach, d["achtemp"], what=f"{case} achtemp alongside")


@pytest.mark.parametrize("shape", [(128, 2, 8192, 409), (7, 3, 100, 9), (65, 0, 300, 70), (1, 1, 1, 1)])
def test_n1_vs_oracle(fk, fko, shape):
    nb, nv, C, G = shape
    print(f"acht_n1_loc shape {shape}", flush=True)
    p = fko.gpp_random(nb, nv, C, G, seed=nb + G)
    _, _, ref = fko.gpp(p, want_n1loc=True)
    ach, _, n1 = _run(fk, p)
    _close(n1, ref, what=f"{shape} acht_n1_loc")
    assert abs(n1.sum() - ach[2]) <= 1e-12 * abs(ach[2])


def test_n1_mask_off(fk, fko):
    p = fko.gpp_random(16, 2, 256, 20, seed=9, wtilde_re=(3.0, 5.0))
    _, _, ref = fko.gpp(p, mask_on=False, want_n1loc=True)
    _, _, n1 = _run(fk, p, mask=False)
    _close(n1, ref, what="mask off")


def test_n1_needs_three_frequencies(fk, fko):
    p = fko.gpp_random(4, 1, 32, 4, seed=1)
    p.wx = p.wx[:2]
    with pytest.raises(fk.FkError) as e:
        _run(fk, p, memory="host")
    assert e.value.code == fk.FK_ERR_ARG and "nfreq" in fk.last_error()


def test_n1_mixed_vs_fp64_oracle(fk, fko):
    torch = _torch()
    p = fko.gpp_random(48, 4, 2048, 102, seed=2, wtilde_re=(3.0, 5.0))
    for k in ("aqsm", "aqsn", "I_eps", "wtilde"):
        setattr(p, k, getattr(p, k).astype(np.complex64).astype(np.complex128))
    _, _, ref = fko.gpp(p, want_n1loc=True)
    t = _dev(p, torch.complex64)
    _, _, _, n1 = fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, t["aqsm"], t["aqsn"], t["I_eps"],
                                t["wtilde"], t["vcoul"], p.inv_igp_index, p.indinv, precision="mixed",
                                acht_n1=True)
    torch.cuda.synchronize()
    # the per-band sum runs in FP64 on the complex64 values: FP64 accuracy
    _close(n1.cpu().numpy(), ref, what="mixed acht_n1_loc")


def test_n1_full_size_sums_to_achtemp(fk, fko):
    torch = _torch()
    nb, nv, C, G = 512, 2, 32768, 1638
    t = {}
    for name, n, re in (("aqsm", nb * C, (-0.5, 0.5)), ("aqsn", nb * C, (-0.5, 0.5)),
                        ("I_eps", G * C, (-0.5, 0.5)), ("wtilde", G * C, (0.5, 2.0))):
        t[name] = torch.empty(n, dtype=torch.complex128, device="cuda:0")
        im = (-0.01, 0.01) if name == "wtilde" else (-0.5, 0.5)
        fk.fill_complex(t[name], 1, fko.STREAM[name], re, im)
    v = torch.empty(C, dtype=torch.float64, device="cuda:0")
    fk.fill_real(v, 1, fko.STREAM["vcoul"], 0.1, 10.0)
    inv, indinv = fko.index_maps(C, G, "fortran")
    wx = fko.wx_array(3)
    ach, _, _, n1 = fk.gpp_kernel(nb, nv, C, G, wx, t["aqsm"], t["aqsn"], t["I_eps"], t["wtilde"], v, inv,
                                  indinv, acht_n1=True)
    torch.cuda.synchronize()
    a2 = complex(ach.cpu()[2])
    s = n1.cpu().numpy().sum()
    assert abs(s - a2) <= 1e-12 * abs(a2)
    # the oracle on bands 0 and 300 at the full ngpown x ncouls shape
    sel = [0, 300]
    p = fko.GppProblem(2, 1, C, G, wx, t["aqsm"].view(nb, C)[sel].cpu().numpy(),
                       t["aqsn"].view(nb, C)[sel].cpu().numpy(), t["I_eps"].view(G, C).cpu().numpy(),
                       t["wtilde"].view(G, C).cpu().numpy(), v.cpu().numpy(), inv, indinv)
    _, _, ref = fko.gpp(p, want_n1loc=True, nthreads=8)
    _close(n1.cpu().numpy()[sel], ref, what="full-size sampled bands")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_n1_in_caller_workspace(fk, fko, precision):
    """A caller workspace of gpp_workspace_size(acht_n1=True) bytes serves the
    per-band output as well (no allocation of its own); one of the smaller
    gpp_workspace_size() leaves that output its per-call allocation.  Both
    give the outputs of a call without a workspace, bit for bit, and a
    workspace filled with NaN bytes changes nothing."""
    torch = _torch()
    nb, nv, C, G = 96, 3, 4096, 205
    p = fko.gpp_random(nb, nv, C, G, seed=33, wtilde_re=(3.0, 5.0))
    cdt = torch.complex64 if precision == "mixed" else torch.complex128
    t = _dev(p, cdt)
    big = fk.gpp_workspace_size(nb, nv, C, G, len(p.wx), precision=precision, acht_n1=True)
    small = fk.gpp_workspace_size(nb, nv, C, G, len(p.wx), precision=precision)
    assert big >= small + 16 * G * C
    print(f"acht_n1 workspace {precision}: {small} B without, {big} B with the per-band output", flush=True)

    def call(ws):
        r = fk.gpp_kernel(nb, nv, C, G, p.wx, t["aqsm"], t["aqsn"], t["I_eps"], t["wtilde"], t["vcoul"],
                          p.inv_igp_index, p.indinv, precision=precision, acht_n1=True, workspace=ws)
        torch.cuda.synchronize()
        return [x.cpu().numpy() for x in (r[0], r[1], r[3])]

    ref = call(None)
    for size in (big, small):
        ws = torch.full((size,), 255, dtype=torch.uint8, device="cuda:0")
        got = call(ws)
        for g, r, what in zip(got, ref, ("achtemp", "asxtemp", "acht_n1_loc")):
            assert np.array_equal(g, r), f"{precision} workspace {size} B: {what}"
    _, _, n1ref = fko.gpp(p, want_n1loc=True)
    if precision == "fp64":
        _close(ref[2], n1ref, what="acht_n1_loc with a caller workspace")

