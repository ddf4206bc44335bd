"""GPU parity of the static-CH remainder achstemp (gpp_kernel_ex), the sum of
the Fortran GPP program gppkernel.deslippe.09.30.2014.f90:192-230 taken
from the GPP kernels' own tiles, against the oracle (itself pinned to the
compiled reference's achsDtemp_Kernel, the same sum), the golden vector and
closed forms.  FP64 bar 1e-12; the int8-split mixed path 1e-8 against the
FP64 oracle on the same complex64 values; the FP32 MFMA path 1e-5.
Asking for achstemp must leave achtemp / asxtemp bit for bit unchanged."""
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


def _run(fk, p, memory="device", cdt="c128", **kw):
    torch = _torch()
    args = ("aqsm", "aqsn", "I_eps", "wtilde")
    if memory == "host":
        np_dt = np.complex128 if cdt == "c128" else np.complex64
        t = {k: np.ascontiguousarray(getattr(p, k).astype(np_dt)) for k in args}
        v = p.vcoul
    else:
        dt = torch.complex128 if cdt == "c128" else torch.complex64
        f = lambda a, d: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0").to(d)  # noqa: E731
        t = {k: f(getattr(p, k), dt) for k in args}
        v = f(p.vcoul, torch.float64)
    out = fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, t["aqsm"], t["aqsn"], t["I_eps"], t["wtilde"], v,
                        p.inv_igp_index, p.indinv, to1=p.to1, limittwo=p.limittwo, limitone=p.limitone, occ=p.occ,
                        **kw)
    if memory != "host":
        torch.cuda.synchronize()
        out = tuple(x.cpu().numpy() if hasattr(x, "cpu") else x for x in out)
    return out


def _achs(out):
    return complex(np.asarray(out[-1]).reshape(-1)[0])


@pytest.mark.parametrize("memory", ["host", "device"])
def test_golden_vector(fk, fko, memory):
    d = np.load(os.path.join(GOLDEN, "gpp_achstemp.npz"))
    p = fko.GppProblem(int(d["nb"]), int(d["nv"]), int(d["ncouls"]), int(d["ngpown"]), d["wx"], d["aqsm"],
                       d["aqsn"], d["I_eps"], d["wtilde"], d["vcoul"], d["inv_igp_index"], d["indinv"])
    ref = complex(d["achstemp"])
    got = _achs(_run(fk, p, memory, achstemp=True))
    print(f"achstemp golden ({memory}): {got} vs {ref}", flush=True)
    assert abs(got - ref) <= RTOL * abs(ref)


@pytest.mark.parametrize("config", [0, 1, 2, 3, 4])
def test_vs_oracle_every_tile_config(fk, fko, config):
    p = fko.gpp_random(128, 2, 8192, 409, seed=0)
    print(f"achstemp 128/2/8192/409 config {config}", flush=True)
    ref = fko.gpp_achstemp(p)
    plain = _run(fk, p, config=config)
    out = _run(fk, p, config=config, achstemp=True)
    assert abs(_achs(out) - ref) <= RTOL * abs(ref), (_achs(out), ref)
    assert np.array_equal(out[0], plain[0]) and np.array_equal(out[1], plain[1])


@pytest.mark.parametrize("shape", [(7, 3, 100, 9), (1, 1, 1, 1), (33, 0, 257, 17), (64, 64, 640, 33)])
def test_ragged_shapes(fk, fko, shape):
    nb, nv, C, G = shape
    p = fko.gpp_random(nb, nv, C, G, seed=sum(shape))
    ref = fko.gpp_achstemp(p)
    got = _achs(_run(fk, p, achstemp=True))
    assert abs(got - ref) <= RTOL * abs(ref), (shape, got, ref)


def test_skipped_columns_and_acht_n1(fk, fko):
    p = fko.gpp_random(20, 3, 300, 30, seed=5)
    ind = p.indinv.copy()
    ind[p.inv_igp_index[4]] = -1
    ind[p.inv_igp_index[9]] = p.ncouls + 5
    p.indinv = ind
    ref = fko.gpp_achstemp(p)
    out = _run(fk, p, achstemp=True, acht_n1=True)
    assert abs(_achs(out) - ref) <= RTOL * abs(ref)
    _, _, n1ref = fko.gpp(p, want_n1loc=True)
    assert np.max(np.abs(out[3] - n1ref)) <= RTOL * np.max(np.abs(n1ref))


def test_full_size_closed_form(fk, fko):
    """512/2/32768/20, constant inputs 0.5+0.5i, vcoul = 1:
    achstemp = -0.5 nb ngpown ncouls (0.25+0.25i)."""
    torch = _torch()
    nb, nv, C, npg = 512, 2, 32768, 20
    G = C // npg
    d = torch.device("cuda:0")
    z = torch.full((nb, C), 0.5 + 0.5j, dtype=torch.complex128, device=d)
    g = torch.full((G, C), 0.5 + 0.5j, dtype=torch.complex128, device=d)
    v = torch.ones(C, dtype=torch.float64, device=d)
    inv, indinv = fko.index_maps(C, G, "cpp")
    wx = fko.wx_array(3)
    out = fk.gpp_kernel(nb, nv, C, G, wx, z, z, g, g, v, inv, indinv, achstemp=True, time_kernel=True)
    torch.cuda.synchronize()
    want = -0.5 * nb * G * C * (0.25 + 0.25j)
    got = complex(out[-1].cpu()[0])
    print(f"achstemp full size: {got} want {want}; kernel {out[2]:.3f} ms", flush=True)
    assert abs(got - want) <= RTOL * abs(want)


@pytest.mark.parametrize("memory", ["device", "host"])
def test_mixed_int8_split(fk, fko, memory):
    p = fko.gpp_random(96, 4, 2048, 102, seed=3)
    for k in ("aqsm", "aqsn", "I_eps", "wtilde"):
        setattr(p, k, getattr(p, k).astype(np.complex64).astype(np.complex128))
    ref = fko.gpp_achstemp(p)
    plain = _run(fk, p, memory, cdt="c64", precision="mixed")
    out = _run(fk, p, memory, cdt="c64", precision="mixed", achstemp=True)
    err = abs(_achs(out) - ref) / abs(ref)
    print(f"achstemp int8 split ({memory}): rel err {err:.2e}", flush=True)
    assert err <= 1e-8
    assert np.array_equal(out[0], plain[0]) and np.array_equal(out[1], plain[1])


def test_mixed_fp32_mfma(fk, fko):
    p = fko.gpp_random(96, 4, 2048, 102, seed=4)
    for k in ("aqsm", "aqsn", "I_eps", "wtilde"):
        setattr(p, k, getattr(p, k).astype(np.complex64).astype(np.complex128))
    ref = fko.gpp_achstemp(p)
    out = _run(fk, p, cdt="c64", precision="mixed", config=2, achstemp=True)
    err = abs(_achs(out) - ref) / abs(ref)
    print(f"achstemp FP32 MFMA: rel err {err:.2e}", flush=True)
    assert err <= 1e-5
