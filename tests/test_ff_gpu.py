"""GPU parity of ff_kernel (through the C ABI) against the CPU oracle, the
golden vectors of the reference's achsDtemp kernel and program, and, at the
full 256/4096/4096/32 size, against a sampled oracle and size-independent
properties.  Bar: relative error <= 1e-12 for FP64."""
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


def _dev(p):
    torch = _torch()
    d = torch.device("cuda:0")
    f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(d).to(dt)  # noqa: E731
    return dict(aqsm=f(p.aqsm, torch.complex128), aqsn=f(p.aqsn, torch.complex128),
                I_eps=f(p.I_eps, torch.complex128), vcoul=f(p.vcoul, torch.float64))


def _run_dev(fk, p, **kw):
    t = _dev(p)
    schDt, schDv, achs, ms = fk.ff_kernel(p.nb, p.ncouls, p.ngpown, p.nfreq, t["aqsm"], t["aqsn"], t["I_eps"],
                                          t["vcoul"], p.inv_igp_index, p.indinv, **kw)
    _torch().cuda.synchronize()
    return schDt.cpu().numpy(), (None if schDv is None else schDv.cpu().numpy()), complex(achs.cpu()[0]), ms


def _check(fko, fk, p, achs_freq=0, **kw):
    rt, rv, ra = fko.ff(p, achs_freq=achs_freq)
    gt, gv, ga, _ = _run_dev(fk, p, achs_freq=achs_freq, **kw)
    assert _rel(gt, rt) <= RTOL, f"schDt rel {_rel(gt, rt):.3e}"
    assert _rel(gv, rv) <= RTOL, f"schDv rel {_rel(gv, rv):.3e}"
    assert abs(ga - ra) <= RTOL * abs(ra)
    return gt, gv, ga


def test_golden_achs_kernel_host_memory(fk, fko):
    """The reference's achsDtemp_Kernel (FFKer_cori/ffKernel.cpp:107-131, frequency
    index 1) on seeded inputs, run from host memory."""
    d = np.load(os.path.join(GOLDEN, "ff_achs.npz"))
    p = fko.FfProblem(int(d["nb"]), int(d["ncouls"]), int(d["ngpown"]), int(d["nfreq"]), d["aqsm"], d["aqsn"],
                      d["I_eps"], d["vcoul"], d["inv_igp_index"], d["indinv"])
    schDt, schDv, achs, _ = fk.ff_problem(p, achs_freq=1)
    ref = complex(d["achsDtemp"])
    assert abs(complex(achs[0]) - ref) <= RTOL * abs(ref)
    rt, rv, _ = fko.ff(p, achs_freq=1)
    assert _rel(schDt, rt) <= RTOL and _rel(schDv, rv) <= RTOL


def test_golden_reference_program(fk, fko):
    """achsDtemp printed by the reference's ffKernel program on its constant
    inputs, and the contraction's closed form G*C*(0.25+0.25i)."""
    torch = _torch()
    d = np.load(os.path.join(GOLDEN, "ff_main.npz"))
    for (nb, nv, C, G, nf, nfe), ref in zip(d["shapes"], d["achsDtemp"]):
        nb, C, G, nf = int(nb), int(C), int(G), int(nf)
        z = torch.full((nb, C), 0.5 + 0.5j, dtype=torch.complex128, device="cuda:0")
        I = torch.full((nf, G, C), 0.5 + 0.5j, dtype=torch.complex128, device="cuda:0")
        v = torch.ones(C, dtype=torch.float64, device="cuda:0")
        inv = np.arange(G, dtype=np.int32)
        indinv = np.arange(C + 1, dtype=np.int32)
        indinv[C] = C - 1
        schDt, _, achs, _ = fk.ff_kernel(nb, C, G, nf, z, z, I, v, inv, indinv, achs_freq=1)
        torch.cuda.synchronize()
        assert abs(complex(achs.cpu()[0]) - ref) <= RTOL * abs(ref)
        assert np.allclose(schDt.cpu().numpy(), G * C * (0.25 + 0.25j), rtol=1e-15, atol=0)


@pytest.mark.parametrize("shape", [(6, 40, 5, 3), (1, 1, 1, 1), (33, 257, 17, 2), (70, 300, 130, 4),
                                   (64, 64, 64, 1), (129, 1000, 200, 3)])
def test_vs_oracle_shapes(fk, fko, shape):
    nb, C, G, nf = shape
    print(f"ff shape {shape}", flush=True)
    p = fko.ff_random(nb, C, G, nf, seed=sum(shape))
    _check(fko, fk, p)


def test_permuted_and_skipped_columns(fk, fko):
    p = fko.ff_random(20, 96, 12, 2, seed=11)
    inv = p.inv_igp_index.copy()
    inv[[1, 3]] = inv[[3, 1]]
    inv[5] = p.ncouls                      # indigp == ncouls -> igp = ncouls - 1
    ind = p.indinv.copy()
    ind[inv[7]] = -1                       # skipped column
    p.inv_igp_index, p.indinv = inv, ind
    _check(fko, fk, p)


@pytest.mark.parametrize("achs_freq", [0, 1, 2])
def test_achs_frequency(fk, fko, achs_freq):
    p = fko.ff_random(9, 80, 10, 3, seed=3)
    _check(fko, fk, p, achs_freq=achs_freq)


@pytest.mark.parametrize("config", [1, 2, 3])
def test_tile_configs(fk, fko, config):
    p = fko.ff_random(100, 333, 150, 2, seed=config)
    _check(fko, fk, p, config=config)


@pytest.mark.parametrize("config", [1, 2])
@pytest.mark.parametrize("shape", [(70, 3072, 40, 2), (20, 2000, 33, 3), (5, 1100, 17, 1)])
def test_ig_split_small_problems(fk, fko, config, shape):
    """Few tiles and a long ig range: the FP64 tiles split the ig range over
    up to 8 workgroups (8, 4 and 4 parts here, the last one shorter), whose
    partials the reduction sums; with skipped and permuted columns."""
    nb, C, G, nf = shape
    print(f"ff ig split config {config} shape {shape}", flush=True)
    p = fko.ff_random(nb, C, G, nf, seed=7 + nb)
    inv = p.inv_igp_index.copy()
    inv[[0, 2]] = inv[[2, 0]]
    ind = p.indinv.copy()
    ind[inv[4]] = -1
    p.inv_igp_index, p.indinv = inv, ind
    _check(fko, fk, p, config=config)


def test_device_stream_workspace_timer_and_no_schDv(fk, fko):
    torch = _torch()
    p = fko.ff_random(40, 200, 50, 3, seed=5)
    rt, _, ra = fko.ff(p)
    n = fk.ff_workspace_size(p.nb, p.ncouls, p.ngpown, p.nfreq)
    ws = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        gt, gv, ga, ms = _run_dev(fk, p, stream=s, workspace=ws, time_kernel=True, want_schDv=False)
    assert gv is None and ms is not None and ms > 0
    assert _rel(gt, rt) <= RTOL and abs(ga - ra) <= RTOL * abs(ra)


# ---------------------------------------------------------------------------
# full size: nb=256, ncouls=ngpown=4096, nFreq=32 (I_eps 8.59 GB on the device)
# ---------------------------------------------------------------------------
NB, C, G, NF = 256, 4096, 4096, 32


@pytest.fixture(scope="module")
def full(fk):
    torch = _torch()
    d = torch.device("cuda:0")
    seed = 4
    aqsm = torch.empty(NB, C, dtype=torch.complex128, device=d)
    aqsn = torch.empty(NB, C, dtype=torch.complex128, device=d)
    I = torch.empty(NF, G, C, dtype=torch.complex128, device=d)
    v = torch.empty(C, dtype=torch.float64, device=d)
    fk.fill_complex(aqsm, seed, 1, (-0.5, 0.5), (-0.5, 0.5))
    fk.fill_complex(aqsn, seed, 2, (-0.5, 0.5), (-0.5, 0.5))
    fk.fill_complex(I, seed, 6, (-0.5, 0.5), (-0.5, 0.5))
    fk.fill_real(v, seed, 5, 0.1, 10.0)
    inv = np.arange(G, dtype=np.int32)
    indinv = np.arange(C + 1, dtype=np.int32)
    indinv[C] = C - 1
    print("ff full-size inputs filled", flush=True)
    out = fk.ff_kernel(NB, C, G, NF, aqsm, aqsn, I, v, inv, indinv)
    torch.cuda.synchronize()
    yield dict(aqsm=aqsm, aqsn=aqsn, I=I, v=v, inv=inv, indinv=indinv, out=out, seed=seed)
    del I
    torch.cuda.empty_cache()


def test_full_size_sampled_oracle(fk, fko, full):
    """Rows of schDt depend only on their own band and columns only on their
    own frequency: the oracle on bands {0, 131, 255} x frequencies {0, 17, 31}
    checks the full-size launch entry by entry."""
    bands, freqs = [0, 131, 255], [0, 17, 31]
    p = fko.FfProblem(len(bands), C, G, len(freqs), full["aqsm"][bands].cpu().numpy(),
                      full["aqsn"][bands].cpu().numpy(), full["I"][freqs].cpu().numpy(),
                      full["v"].cpu().numpy(), full["inv"], full["indinv"])
    rt, rv, _ = fko.ff(p, nthreads=8)
    schDt, schDv, _, _ = full["out"]
    gt = schDt.cpu().numpy()[np.ix_(bands, freqs)]
    gv = schDv.cpu().numpy()[np.ix_(bands, freqs)]
    assert _rel(gt, rt) <= RTOL and _rel(gv, rv) <= RTOL


def test_full_size_achs_is_sum_of_schDv(full):
    schDt, schDv, achs, _ = full["out"]
    a = complex(achs.cpu()[0])
    s = -0.5 * schDv[:, 0].cpu().numpy().sum()
    assert abs(a - s) <= 1e-13 * abs(s)


def test_full_size_deterministic_and_linear(fk, full):
    torch = _torch()
    schDt, _, _, _ = full["out"]
    again = fk.ff_kernel(NB, C, G, NF, full["aqsm"], full["aqsn"], full["I"], full["v"], full["inv"],
                         full["indinv"], want_schDv=False)[0]
    torch.cuda.synchronize()
    assert torch.equal(again, schDt)
    # schDt is linear in aqsn: doubling it doubles the result exactly
    twice = fk.ff_kernel(NB, C, G, NF, full["aqsm"], full["aqsn"] * 2, full["I"], full["v"], full["inv"],
                         full["indinv"], want_schDv=False)[0]
    torch.cuda.synchronize()
    assert torch.equal(twice, 2 * schDt)
