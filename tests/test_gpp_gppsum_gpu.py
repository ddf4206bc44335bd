"""GPU parity of the gppsum = 1 form of gpp_kernel (FK_GPP_GPPSUM, the band
loop of gppKerSeq.cpp:76-261,383-472) through the C ABI: the golden vectors
that the reference's compiled flagOCC_solver / noflagOCC_solver wrote, the
oracle (pinned by those vectors) on ragged and edge shapes, and at the full
512/2/32768/20 size the oracle on the two occupied bands plus bitwise
determinism and exact linearity.  Bar: 1e-12 relative."""
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
    for g, r in zip(got, ref):
        assert abs(g - r) <= rtol * abs(r) + 1e-3 * rtol * scale, f"{what}: got {g} want {r}"


def _dev(p):
    torch = _torch()
    f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0").to(dt)  # noqa: E731
    return dict(aqsm=f(p.aqsm, torch.complex128), aqsn=f(p.aqsn, torch.complex128),
                I_eps=f(p.I_eps, torch.complex128), wtilde=f(p.wtilde, torch.complex128),
                vcoul=f(p.vcoul, torch.float64))


def _run(fk, p, memory="device", **kw):
    if memory == "host":
        ach, asx, ms = fk.gpp_problem(p, semantics=fk.FK_GPP_GPPSUM, **kw)
        return ach, asx, ms
    t = _dev(p)
    ach, asx, ms = fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, t["aqsm"], t["aqsn"], t["I_eps"],
                                 t["wtilde"], t["vcoul"], p.inv_igp_index, p.indinv, to1=p.to1,
                                 limittwo=p.limittwo, limitone=p.limitone, occ=p.occ,
                                 semantics=fk.FK_GPP_GPPSUM, **kw)
    _torch().cuda.synchronize()
    return ach.cpu().numpy(), asx.cpu().numpy(), ms


def _from_npz(fko, d):
    return fko.GppProblem(int(d["nb"]), int(d["nv"]), int(d["ncouls"]), int(d["ngpown"]), d["wx"], d["aqsm"],
                          d["aqsn"], d["I_eps"], d["wtilde"], d["vcoul"], d["inv_igp_index"], d["indinv"],
                          occ=float(d["occ"]))


@pytest.mark.parametrize("case", ["random", "blocks", "break"])
@pytest.mark.parametrize("memory", ["device", "host"])
def test_gppsum_golden(fk, fko, case, memory):
    d = np.load(os.path.join(GOLDEN, f"gppsum_{case}.npz"))
    p = _from_npz(fko, d)
    print(f"gppsum golden {case} ({memory}): {p.nb}/{p.nv}/{p.ncouls}/{p.ngpown} nw {len(p.wx)}", flush=True)
    ach, asx, _ = _run(fk, p, memory)
    _close(ach, d["achtemp"], what=f"{case} achtemp")
    _close(asx, d["asxtemp"], what=f"{case} asxtemp")


# (nb, nv, ncouls, ngpown, nw, wtilde.re range, column map)
SHAPES = [(1, 1, 64, 3, 3, (2.5, 5.5), "fortran"), (7, 0, 600, 6, 1, (2.5, 5.5), "fortran"),
          (9, 9, 700, 7, 5, (0.5, 2.0), "cpp"), (33, 4, 1537, 29, 16, (2.5, 5.5), "fortran"),
          (64, 2, 4096, 205, 3, (2.5, 5.5), "cpp"), (130, 16, 2048, 100, 11, (0.5, 5.5), "fortran")]


@pytest.mark.parametrize("shape", SHAPES, ids=[f"{s[0]}x{s[1]}x{s[2]}x{s[3]}x{s[4]}" for s in SHAPES])
def test_gppsum_vs_oracle(fk, fko, shape):
    nb, nv, C, G, nw, wre, mk = shape
    p = fko.gpp_random(nb, nv, C, G, nw=nw, seed=60 + nb, wtilde_re=wre, map_kind=mk)
    print(f"gppsum oracle {nb}/{nv}/{C}/{G} nw {nw}", flush=True)
    ra, rs, _ = fko.gpp_gppsum(p)
    ach, asx, _ = _run(fk, p)
    _close(ach, ra, what="achtemp")
    if nv:
        _close(asx, rs, what="asxtemp")
    else:
        assert np.all(asx == 0)


def test_gppsum_skipped_columns_and_occupied_stop(fk, fko):
    p = fko.gpp_random(12, 5, 1300, 13, nw=4, seed=71, wtilde_re=(2.5, 5.5))
    ind = p.indinv.copy()
    ind[p.inv_igp_index[3]] = -1
    ind[p.inv_igp_index[9]] = -5
    p.indinv = ind
    ie = p.I_eps.copy()
    for m, ig in ((1, 3), (4, 200), (6, 90), (12, 0)):
        ie[m, ig] = 2e-9 - 1e-9j         # the occupied loop stops at these points
    p.I_eps = ie
    ra, rs, _ = fko.gpp_gppsum(p)
    ach, asx, _ = _run(fk, p)
    _close(ach, ra, what="achtemp")
    _close(asx, rs, what="asxtemp")


def test_gppsum_kernel_timing_and_no_band_problem(fk, fko):
    p = fko.gpp_random(16, 2, 1024, 51, seed=72, wtilde_re=(2.5, 5.5))
    ach, asx, ms = _run(fk, p, time_kernel=True)
    assert ms is not None and ms > 0
    torch = _torch()
    inv, indinv = fko.index_maps(256, 12, "fortran")
    one = torch.zeros(1, dtype=torch.complex128, device="cuda:0")
    w = torch.ones(12 * 256, dtype=torch.complex128, device="cuda:0")
    ach0, asx0, _ = fk.gpp_kernel(0, 0, 256, 12, fko.wx_array(3), one, one, w, w,
                                  torch.ones(256, dtype=torch.float64, device="cuda:0"), inv, indinv,
                                  semantics=fk.FK_GPP_GPPSUM)
    torch.cuda.synchronize()
    assert torch.count_nonzero(ach0) == 0 and torch.count_nonzero(asx0) == 0


FULL = (512, 2, 32768, 1638)


@pytest.fixture(scope="module")
def full_random(fk, fko):
    torch = _torch()
    nb, nv, C, G = FULL
    t = {}
    for name, rows, re in (("aqsm", nb, (-0.5, 0.5)), ("aqsn", nb, (-0.5, 0.5)), ("I_eps", G, (-0.5, 0.5)),
                           ("wtilde", G, (2.5, 5.5))):
        t[name] = torch.empty((rows, C), dtype=torch.complex128, device="cuda:0")
        im = (-0.01, 0.01) if name == "wtilde" else (-0.5, 0.5)
        fk.fill_complex(t[name], 7, fko.STREAM[name], re, im)
    t["vcoul"] = torch.empty(C, dtype=torch.float64, device="cuda:0")
    fk.fill_real(t["vcoul"], 7, fko.STREAM["vcoul"], 0.1, 10.0)
    torch.cuda.synchronize()
    inv, indinv = fko.index_maps(C, G, "fortran")
    return t, inv, indinv, fko.wx_array(3)


def _full(fk, t, inv, indinv, wx, nb=None, nv=None, vscale=1.0, **kw):
    nb_all, nv_all, C, G = FULL
    nb = nb_all if nb is None else nb
    nv = nv_all if nv is None else nv
    ach, asx, ms = fk.gpp_kernel(nb, nv, C, G, wx, t["aqsm"][:nb], t["aqsn"][:nb], t["I_eps"], t["wtilde"],
                                 t["vcoul"] * vscale, inv, indinv, semantics=fk.FK_GPP_GPPSUM, **kw)
    _torch().cuda.synchronize()
    return ach.cpu().numpy(), asx.cpu().numpy(), ms


def test_gppsum_full_size_deterministic_and_linear(fk, full_random):
    t, inv, indinv, wx = full_random
    a1, s1, ms = _full(fk, t, inv, indinv, wx, time_kernel=True)
    print(f"gppsum 512/2/32768/20: tile kernel {ms:.3f} ms", flush=True)
    a2, s2, _ = _full(fk, t, inv, indinv, wx)
    assert np.array_equal(a1, a2) and np.array_equal(s1, s2)
    a3, s3, _ = _full(fk, t, inv, indinv, wx, vscale=2.0)
    assert np.array_equal(a3, 2 * a1) and np.array_equal(s3, 2 * s1)


def test_gppsum_full_shape_vs_oracle_two_bands(fk, fko, full_random):
    t, inv, indinv, wx = full_random
    nb_all, nv, C, G = FULL
    p = fko.GppProblem(2, 2, C, G, wx, t["aqsm"][:2].cpu().numpy(), t["aqsn"][:2].cpu().numpy(),
                       t["I_eps"].cpu().numpy(), t["wtilde"].cpu().numpy(), t["vcoul"].cpu().numpy(), inv, indinv)
    ra, rs, _ = fko.gpp_gppsum(p)
    a, s, _ = _full(fk, t, inv, indinv, wx, nb=2, nv=2)
    _close(a, ra, what="full shape 2 occupied bands achtemp")
    _close(s, rs, what="full shape 2 occupied bands asxtemp")
    q = fko.GppProblem(3, 0, C, G, wx, t["aqsm"][2:5].cpu().numpy(), t["aqsn"][2:5].cpu().numpy(),
                       p.I_eps, p.wtilde, p.vcoul, inv, indinv)
    ra, _, _ = fko.gpp_gppsum(q)
    ach, _, _ = fk.gpp_kernel(3, 0, C, G, wx, t["aqsm"][2:5], t["aqsn"][2:5], t["I_eps"], t["wtilde"], t["vcoul"],
                              inv, indinv, semantics=fk.FK_GPP_GPPSUM)
    _torch().cuda.synchronize()
    _close(ach.cpu().numpy(), ra, what="full shape 3 unoccupied bands achtemp")
