"""GPU parity of mixed-precision GPP batches on the int8-split contraction
(gpp_batch_create with FK_PREC_MIXED and config 0): the operand-slice, int8
GEMM and FP64 pole kernels of gpp_kernel's mixed path, each launched once
over every problem's blocks, then one reduction, captured into a graph.

Each problem's outputs are held to gpp_kernel's mixed result on the same
complex64 values (the same slices, GEMM tiles and per-element epilogue; the
batch evaluates four frequencies per epilogue pass, so 1e-14 rather than
bitwise), and to the FP64 oracle on the same values at 1e-8, the bar of
gpp_kernel's mixed path.  The per-problem semantics are those of the band
loop BGW-Kernels/gppKer.cpp:303-439."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIXED_RTOL = 1e-8

# (nb, nv, ncouls, ngpown, nfreq): ragged shapes, no occupied band, one band,
# 16 (the most) frequencies, a column count that is no multiple of 64
SHAPES = [(40, 2, 512, 25, 3), (7, 3, 100, 9, 5), (64, 0, 300, 30, 4), (33, 16, 256, 12, 11), (1, 1, 64, 3, 3),
          (96, 8, 1024, 51, 7), (20, 4, 200, 10, 16), (130, 5, 1500, 75, 1)]


def _torch():
    import torch
    assert torch.cuda.is_available(), "no ROCm device visible"
    return torch


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _problems(fko, shapes, seed0=30):
    """complex64 device problems and their oracle problems on the same
    (complex64-rounded) values"""
    torch = _torch()
    probs, oracle = [], []
    for k, (nb, nv, C, G, nw) in enumerate(shapes):
        p = fko.gpp_random(nb, nv, C, G, nw=nw, seed=seed0 + k, wtilde_re=(0.5, 2.0) if k % 2 else (3.0, 5.0))
        for key in ("aqsm", "aqsn", "I_eps", "wtilde"):
            setattr(p, key, getattr(p, key).astype(np.complex64).astype(np.complex128))
        f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0").to(dt)  # noqa: E731
        probs.append(dict(nb=nb, nv=nv, ncouls=C, ngpown=G, wx=p.wx, aqsm=f(p.aqsm, torch.complex64),
                          aqsn=f(p.aqsn, torch.complex64), I_eps=f(p.I_eps, torch.complex64),
                          wtilde=f(p.wtilde, torch.complex64), vcoul=f(p.vcoul, torch.float64),
                          inv_igp_index=p.inv_igp_index, indinv=p.indinv, occ=p.occ))
        oracle.append(p)
    return probs, oracle


def _mixed_single(fk, p):
    return fk.gpp_kernel(p["nb"], p["nv"], p["ncouls"], p["ngpown"], p["wx"], p["aqsm"], p["aqsn"], p["I_eps"],
                         p["wtilde"], p["vcoul"], p["inv_igp_index"], p["indinv"], precision="mixed", config=0)


def test_int8_batch_vs_single_calls_and_oracle(fk, fko):
    torch = _torch()
    probs, oracle = _problems(fko, SHAPES)
    print(f"int8-split batch of {len(probs)} problems: create", flush=True)
    b = fk.GppBatch(probs, precision="mixed", config=0)
    assert b.config == 0
    ach, asx, _ = b.run()
    torch.cuda.synchronize()
    print(f"int8-split batch: workspace {b.workspace_bytes} B", flush=True)
    worst_single, worst_oracle = 0.0, 0.0
    for k, p in enumerate(oracle):
        a1, s1, _ = _mixed_single(fk, probs[k])
        torch.cuda.synchronize()
        ga, gs = ach[k].cpu().numpy(), asx[k].cpu().numpy()
        ea = _rel(ga, a1.cpu().numpy())
        es = _rel(gs, s1.cpu().numpy()) if p.nv else 0.0
        worst_single = max(worst_single, ea, es)
        assert ea <= 1e-14 and es <= 1e-14, f"problem {k} vs gpp_kernel: {ea:.2e} {es:.2e}"
        ra, rs, _ = fko.gpp(p)
        worst_oracle = max(worst_oracle, _rel(ga, ra))
        if p.nv:
            worst_oracle = max(worst_oracle, _rel(gs, rs))
        else:
            assert np.all(gs == 0)
    print(f"int8-split batch: worst vs gpp_kernel mixed {worst_single:.2e}, vs the oracle {worst_oracle:.2e}",
          flush=True)
    assert worst_oracle <= MIXED_RTOL
    b.close()


def test_int8_batch_replays_see_new_values_and_no_band_problem(fk, fko):
    """Replays read the inputs' current values (vcoul doubled before each
    replay doubles every output exactly: the slices' scales absorb a power of
    two), runs are deterministic, and a problem without bands gives zeros."""
    torch = _torch()
    probs, _ = _problems(fko, SHAPES[:3], seed0=40)
    probs.append(dict(nb=0, nv=0, ncouls=128, ngpown=6, wx=np.array([3.0, 4.0, 5.0, 6.0]),
                      aqsm=torch.zeros(1, dtype=torch.complex64, device="cuda"),
                      aqsn=torch.zeros(1, dtype=torch.complex64, device="cuda"),
                      I_eps=torch.zeros(6 * 128, dtype=torch.complex64, device="cuda"),
                      wtilde=torch.ones(6 * 128, dtype=torch.complex64, device="cuda"),
                      vcoul=torch.ones(128, dtype=torch.float64, device="cuda"),
                      inv_igp_index=np.arange(6, dtype=np.int32) * 20, indinv=np.arange(129, dtype=np.int32)))
    b = fk.GppBatch(probs, precision="mixed", config=0)
    base_a, base_s, _ = b.run()
    base_a = [t.clone() for t in base_a]
    base_s = [t.clone() for t in base_s]
    assert torch.count_nonzero(base_a[3]) == 0 and torch.count_nonzero(base_s[3]) == 0
    again_a, again_s, _ = b.run()
    torch.cuda.synchronize()
    for k in range(4):
        assert torch.equal(again_a[k], base_a[k]) and torch.equal(again_s[k], base_s[k]), f"rerun problem {k}"
    for rep in range(3):
        for p in probs[:3]:
            p["vcoul"].mul_(2.0)
        a, s, ms = b.run(time_it=True)
        torch.cuda.synchronize()
        print(f"replay {rep}: {ms:.3f} ms", flush=True)
        assert ms is not None and ms > 0
        for k in range(3):
            assert torch.equal(a[k], base_a[k] * 2.0 ** (rep + 1)), f"replay {rep} problem {k}"
            assert torch.equal(s[k], base_s[k] * 2.0 ** (rep + 1)), f"replay {rep} problem {k}"
        assert torch.count_nonzero(a[3]) == 0
        _ = torch.ones(1 << 20, device="cuda").sum()   # eager work between replays
    b.close()


def test_int8_batch_one_call(fk, fko):
    """gpp_kernel_batched with FK_PREC_MIXED and config 0 runs the same
    contraction: equal to gpp_kernel's mixed result within 1e-14."""
    torch = _torch()
    probs, _ = _problems(fko, SHAPES[:4], seed0=50)
    b = fk.GppBatch(probs, precision="mixed", config=0)     # keeps the structs alive
    n = len(probs)
    arr = (ctypes.POINTER(fk.GppProblem) * n)(*[ctypes.pointer(s) for s in b._structs])
    ach = [torch.empty(len(p["wx"]), dtype=torch.complex128, device="cuda") for p in probs]
    asx = [torch.empty(len(p["wx"]), dtype=torch.complex128, device="cuda") for p in probs]
    pa = (ctypes.c_void_p * n)(*[t.data_ptr() for t in ach])
    px = (ctypes.c_void_p * n)(*[t.data_ptr() for t in asx])
    opt = fk.default_options()
    opt.memory = fk.FK_MEM_DEVICE
    opt.precision = fk.FK_PREC_MIXED
    print(f"gpp_kernel_batched mixed, {n} problems", flush=True)
    assert fk.lib.gpp_kernel_batched(n, arr, ctypes.byref(opt), pa, px) == 0, fk.last_error()
    for k, p in enumerate(probs):
        a1, s1, _ = _mixed_single(fk, p)
        torch.cuda.synchronize()
        assert _rel(ach[k].cpu().numpy(), a1.cpu().numpy()) <= 1e-14, f"problem {k}"
        if p["nv"]:
            assert _rel(asx[k].cpu().numpy(), s1.cpu().numpy()) <= 1e-14, f"problem {k}"
    b.close()
