"""GPU parity of the batched GPP path (one persistent launch for many
problems, captured into a graph) against the CPU oracle and against
gpp_kernel on each problem.  Bar: 1e-12 relative (FP64)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-12

# (nb, nv, ncouls, ngpown, nfreq): uneven shapes, no occupied band, 11 and
# 16 (the most) frequencies
SHAPES = [(40, 2, 512, 25, 3), (7, 3, 100, 9, 5), (64, 0, 300, 30, 4), (33, 16, 256, 12, 11), (1, 1, 64, 3, 3),
          (96, 8, 1024, 51, 7), (20, 4, 200, 10, 16)]


def _torch():
    import torch
    assert torch.cuda.is_available(), "no ROCm device visible"
    return torch


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _problems(fko, shapes, seed0=10, cdt=None, round64=False):
    torch = _torch()
    cdt = cdt or torch.complex128
    probs, oracle = [], []
    for k, (nb, nv, C, G, nw) in enumerate(shapes):
        p = fko.gpp_random(nb, nv, C, G, nw=nw, seed=seed0 + k, wtilde_re=(0.5, 2.0) if k % 2 else (3.0, 5.0))
        if round64:
            for key in ("aqsm", "aqsn", "I_eps", "wtilde"):
                setattr(p, key, getattr(p, key).astype(np.complex64).astype(np.complex128))
        f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0").to(dt)  # noqa: E731
        probs.append(dict(nb=nb, nv=nv, ncouls=C, ngpown=G, wx=p.wx, aqsm=f(p.aqsm, cdt), aqsn=f(p.aqsn, cdt),
                          I_eps=f(p.I_eps, cdt), wtilde=f(p.wtilde, cdt), vcoul=f(p.vcoul, torch.float64),
                          inv_igp_index=p.inv_igp_index, indinv=p.indinv, occ=p.occ))
        oracle.append(p)
    return probs, oracle


@pytest.mark.parametrize("config", [0, 1, 2, 3, 4])
def test_batch_vs_oracle(fk, fko, config):
    torch = _torch()
    probs, oracle = _problems(fko, SHAPES)
    b = fk.GppBatch(probs, config=config)
    ach, asx, _ = b.run()
    torch.cuda.synchronize()
    print(f"batch config requested {config}, used {b.config}, workspace {b.workspace_bytes} B", flush=True)
    assert 1 <= b.config <= 4 and (config == 0 or b.config == config)
    for k, p in enumerate(oracle):
        ra, rs, _ = fko.gpp(p)
        assert _rel(ach[k].cpu().numpy(), ra) <= RTOL, f"problem {k} achtemp"
        if p.nv:
            assert _rel(asx[k].cpu().numpy(), rs) <= RTOL, f"problem {k} asxtemp"
        else:
            assert np.all(asx[k].cpu().numpy() == 0)
    b.close()


def test_batch_equals_single_problem_calls(fk, fko):
    """Same tile configuration -> the same tiles and the same fixed-order sums
    as gpp_kernel, except that batches run 4 frequencies per epilogue pass
    where gpp_kernel runs up to 3 in one pass."""
    torch = _torch()
    probs, _ = _problems(fko, SHAPES[:3])
    b = fk.GppBatch(probs, config=2)
    ach, asx, _ = b.run()
    for k, p in enumerate(probs):
        a1, s1, _ = fk.gpp_kernel(p["nb"], p["nv"], p["ncouls"], p["ngpown"], p["wx"], p["aqsm"], p["aqsn"],
                                  p["I_eps"], p["wtilde"], p["vcoul"], p["inv_igp_index"], p["indinv"], config=2)
        torch.cuda.synchronize()
        assert _rel(ach[k].cpu().numpy(), a1.cpu().numpy()) <= 1e-14
        assert _rel(asx[k].cpu().numpy(), s1.cpu().numpy()) <= 1e-14 or p["nv"] == 0
    b.close()


def test_batch_graph_replays_read_current_inputs(fk, fko):
    """Replays of the captured graph, with eager work between them, each
    see the inputs' current values: doubling vcoul doubles every output."""
    torch = _torch()
    probs, _ = _problems(fko, SHAPES[:4])
    b = fk.GppBatch(probs, config=2)
    base_a, base_s, _ = b.run()
    base_a = [t.clone() for t in base_a]
    base_s = [t.clone() for t in base_s]
    for rep in range(4):
        scale = 2.0 ** (rep % 2)
        for p in probs:
            p["vcoul"].mul_(scale)
        a, s, ms = b.run(time_it=True)
        torch.cuda.synchronize()
        assert ms is not None and ms > 0
        # after replay r the cumulative scale is 2^(number of odd r so far)
        cum = 2.0 ** ((rep + 1) // 2)
        for k in range(len(probs)):
            assert torch.equal(a[k], base_a[k] * cum), f"replay {rep} problem {k}"
            assert torch.equal(s[k], base_s[k] * cum), f"replay {rep} problem {k}"
        _ = torch.ones(1 << 20, device="cuda").sum()   # eager work between replays
    b.close()


MIXED_RTOL = 1e-8      # the int8-split contraction, as gpp_kernel's mixed path
INT8_BATCH_OFF = ("first written for a multi-stream graph form of the int8-split batch, whose first MI355X run "
                  "ended in a host segmentation fault inside gpp_batch_create with the cause not found; that "
                  "form is removed, and test_gpp_batch_int8_gpu.py tests the batched-kernel form that replaced it")


def _mixed_single(fk, p):
    return fk.gpp_kernel(p["nb"], p["nv"], p["ncouls"], p["ngpown"], p["wx"], p["aqsm"], p["aqsn"], p["I_eps"],
                         p["wtilde"], p["vcoul"], p["inv_igp_index"], p["indinv"], precision="mixed")


@pytest.mark.skip(reason=INT8_BATCH_OFF)
def test_batch_mixed_int8_split(fk, fko):
    """Mixed batches with config 0 run the int8-split contraction: each
    problem's outputs equal gpp_kernel's mixed result bit for bit (the same
    kernels on the same values), and are within 1e-8 of the FP64 oracle on
    the same complex64 values."""
    torch = _torch()
    probs, oracle = _problems(fko, SHAPES, cdt=torch.complex64, round64=True)
    b = fk.GppBatch(probs, precision="mixed", config=0)
    ach, asx, _ = b.run()
    torch.cuda.synchronize()
    print(f"int8-split batch: config {b.config}, workspace {b.workspace_bytes} B", flush=True)
    assert b.config == 0
    worst = 0.0
    for k, p in enumerate(oracle):
        a1, s1, _ = _mixed_single(fk, probs[k])
        torch.cuda.synchronize()
        assert torch.equal(ach[k], a1) and torch.equal(asx[k], s1), f"problem {k}"
        ra, rs, _ = fko.gpp(p)
        worst = max(worst, _rel(ach[k].cpu().numpy(), ra))
        if p.nv:
            worst = max(worst, _rel(asx[k].cpu().numpy(), rs))
        else:
            assert np.all(asx[k].cpu().numpy() == 0)
    print(f"int8-split batch worst relative error {worst:.2e}", flush=True)
    assert worst <= MIXED_RTOL
    b.close()


@pytest.mark.skip(reason=INT8_BATCH_OFF)
def test_batch_mixed_int8_split_replays_and_no_band_problem(fk, fko):
    """Replays of the int8-split graph see the inputs' current values, and a
    problem without bands gives zeros."""
    torch = _torch()
    probs, _ = _problems(fko, SHAPES[:3], cdt=torch.complex64, round64=True)
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
    for rep in range(3):
        for p in probs[:3]:
            p["vcoul"].mul_(2.0)
        a, s, ms = b.run(time_it=True)
        torch.cuda.synchronize()
        assert ms is not None and ms > 0
        for k in range(3):
            assert torch.equal(a[k], base_a[k] * 2.0 ** (rep + 1)), f"replay {rep} problem {k}"
            assert torch.equal(s[k], base_s[k] * 2.0 ** (rep + 1)), f"replay {rep} problem {k}"
        _ = torch.ones(1 << 20, device="cuda").sum()   # eager work between replays
    b.close()


@pytest.mark.parametrize("config", [1, 2, 4])
def test_batch_mixed_fp32_tiles(fk, fko, config):
    """Mixed batches with config 1..4 use the FP32 MFMA tile kernel; its error
    against the FP64 oracle on the same complex64 values is reported and
    bounded."""
    torch = _torch()
    probs, oracle = _problems(fko, SHAPES, cdt=torch.complex64, round64=True)
    b = fk.GppBatch(probs, precision="mixed", config=config)
    ach, asx, _ = b.run()
    torch.cuda.synchronize()
    assert b.config == config
    worst = 0.0
    for k, p in enumerate(oracle):
        ra, rs, _ = fko.gpp(p)
        worst = max(worst, _rel(ach[k].cpu().numpy(), ra))
        if p.nv:
            worst = max(worst, _rel(asx[k].cpu().numpy(), rs))
    print(f"FP32 tile batch worst relative error {worst:.2e} (config {b.config})", flush=True)
    assert worst < 1e-5
    b.close()


def test_kernel_batched_one_call(fk, fko):
    import ctypes
    torch = _torch()
    probs, oracle = _problems(fko, SHAPES[:2])
    structs, keep = [], []
    for p in probs:
        wx = np.ascontiguousarray(p["wx"]); inv = np.ascontiguousarray(p["inv_igp_index"], np.int32)
        ind = np.ascontiguousarray(p["indinv"], np.int32)
        keep += [wx, inv, ind]
        structs.append(fk.GppProblem(p["nb"], p["nv"], p["ncouls"], p["ngpown"], len(wx), wx.ctypes.data,
                                     p["aqsm"].data_ptr(), p["aqsn"].data_ptr(), p["I_eps"].data_ptr(),
                                     p["wtilde"].data_ptr(), p["vcoul"].data_ptr(), inv.ctypes.data, ind.ctypes.data,
                                     1e-6, 0.25, 2.5e5, 1.0))
    arr = (ctypes.POINTER(fk.GppProblem) * 2)(*[ctypes.pointer(s) for s in structs])
    ach = [torch.empty(len(p["wx"]), dtype=torch.complex128, device="cuda") for p in probs]
    asx = [torch.empty(len(p["wx"]), dtype=torch.complex128, device="cuda") for p in probs]
    pa = (ctypes.c_void_p * 2)(*[t.data_ptr() for t in ach])
    px = (ctypes.c_void_p * 2)(*[t.data_ptr() for t in asx])
    opt = fk.default_options()
    opt.memory = fk.FK_MEM_DEVICE
    assert fk.lib.gpp_kernel_batched(2, arr, ctypes.byref(opt), pa, px) == 0, fk.last_error()
    for k, p in enumerate(oracle):
        ra, rs, _ = fko.gpp(p)
        assert _rel(ach[k].cpu().numpy(), ra) <= RTOL and _rel(asx[k].cpu().numpy(), rs) <= RTOL


@pytest.mark.skip(reason=INT8_BATCH_OFF)
def test_kernel_batched_one_call_mixed(fk, fko):
    """gpp_kernel_batched with FK_PREC_MIXED and config 0: the int8-split
    contraction, equal to gpp_kernel's mixed result."""
    import ctypes
    torch = _torch()
    probs, _ = _problems(fko, SHAPES[:3], cdt=torch.complex64, round64=True)
    b = fk.GppBatch(probs, precision="mixed", config=0)     # keeps the structs alive
    arr = (ctypes.POINTER(fk.GppProblem) * 3)(*[ctypes.pointer(s) for s in b._structs])
    ach = [torch.empty(len(p["wx"]), dtype=torch.complex128, device="cuda") for p in probs]
    asx = [torch.empty(len(p["wx"]), dtype=torch.complex128, device="cuda") for p in probs]
    pa = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in ach])
    px = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in asx])
    opt = fk.default_options()
    opt.memory = fk.FK_MEM_DEVICE
    opt.precision = fk.FK_PREC_MIXED
    assert fk.lib.gpp_kernel_batched(3, arr, ctypes.byref(opt), pa, px) == 0, fk.last_error()
    for k, p in enumerate(probs):
        a1, s1, _ = _mixed_single(fk, p)
        torch.cuda.synchronize()
        assert torch.equal(ach[k], a1) and torch.equal(asx[k], s1), f"problem {k}"
    b.close()
