"""gpp_kernel_noalloc: gpp_kernel with its workspace in the library's own
device array.  Through the C ABI, two calls in a row (the second on a
smaller problem, so the array holds the first call's contents) must give
exactly gpp_kernel's results, which the oracle pins; the smaller problem is
also held to the oracle at 1e-12.  The rejections are host-side checks made
before any device work, so they run without a GPU."""
import ctypes

import numpy as np
import pytest

RTOL = 1e-12


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _struct(fk, p, ptrs):
    wx = np.ascontiguousarray(p.wx, np.float64)
    inv = np.ascontiguousarray(p.inv_igp_index, np.int32)
    ind = np.ascontiguousarray(p.indinv, np.int32)
    s = fk.GppProblem(p.nb, p.nv, p.ncouls, p.ngpown, len(wx), wx.ctypes.data, ptrs["aqsm"], ptrs["aqsn"],
                      ptrs["I_eps"], ptrs["wtilde"], ptrs["vcoul"], inv.ctypes.data, ind.ctypes.data,
                      p.to1, p.limittwo, p.limitone, p.occ)
    return s, (wx, inv, ind)


def _noalloc(fk, p, t, precision, config=0):
    """gpp_kernel_noalloc through ctypes on device tensors; returns numpy outputs."""
    import torch
    s, keep = _struct(fk, p, {k: v.data_ptr() for k, v in t.items()})
    opt = fk.default_options()
    opt.memory = fk.FK_MEM_DEVICE
    opt.precision = precision
    opt.config = config
    ach = torch.full((len(p.wx),), 7.0, dtype=torch.complex128, device="cuda:0")
    asx = torch.full((len(p.wx),), 7.0, dtype=torch.complex128, device="cuda:0")
    rc = fk.lib.gpp_kernel_noalloc(ctypes.byref(s), ctypes.byref(opt), ach.data_ptr(), asx.data_ptr(), None)
    assert rc == fk.FK_OK, fk.last_error()
    torch.cuda.synchronize()
    del keep
    return ach.cpu().numpy(), asx.cpu().numpy()


def _device(p, cdt):
    import torch
    f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0").to(dt)  # noqa: E731
    return dict(aqsm=f(p.aqsm, cdt), aqsn=f(p.aqsn, cdt), I_eps=f(p.I_eps, cdt), wtilde=f(p.wtilde, cdt),
                vcoul=f(p.vcoul, torch.float64))


def _reference_call(fk, p, t, precision, config=0):
    import torch
    ach, asx, _ = fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, t["aqsm"], t["aqsn"], t["I_eps"],
                                t["wtilde"], t["vcoul"], p.inv_igp_index, p.indinv, to1=p.to1,
                                limittwo=p.limittwo, limitone=p.limitone, occ=p.occ,
                                precision="mixed" if precision == 1 else "fp64", config=config)
    torch.cuda.synchronize()
    return ach.cpu().numpy(), asx.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,config", [(0, 0), (0, 3), (1, 0), (1, 2)])
def test_noalloc_equals_gpp_kernel_two_calls(fk, fko, precision, config):
    import torch
    assert torch.cuda.is_available(), "no ROCm device visible"
    cdt = torch.complex64 if precision == 1 else torch.complex128
    big = fko.gpp_random(128, 2, 4096, 205, seed=21)   # fits the array in every form (int8 split: 22.6 MiB)
    small = fko.gpp_random(40, 3, 1024, 51, nw=5, seed=22, wtilde_re=(3.0, 5.0))
    if precision == 1:
        for p in (big, small):
            for key in ("aqsm", "aqsn", "I_eps", "wtilde"):
                setattr(p, key, getattr(p, key).astype(np.complex64).astype(np.complex128))
    for p in (big, small):
        print(f"noalloc precision {precision} config {config}: {p.nb}/{p.nv}/{p.ncouls}/{p.ngpown} "
              f"nw {len(p.wx)}", flush=True)
        t = _device(p, cdt)
        got = _noalloc(fk, p, t, precision, config)
        want = _reference_call(fk, p, t, precision, config)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    ra, rs, _ = fko.gpp(small)
    bar = RTOL if precision == 0 else 1e-8 if config == 0 else 1e-5
    assert _rel(got[0], ra) <= bar and _rel(got[1], rs) <= bar


@pytest.mark.gpu
def test_noalloc_repeated_calls_are_bitwise_stable(fk, fko):
    """The array keeps the previous call's contents; a call on another
    problem in between changes nothing."""
    import torch
    a = fko.gpp_random(64, 4, 2048, 102, seed=31)
    b = fko.gpp_random(96, 2, 4096, 204, nw=7, seed=32)
    ta, tb = _device(a, torch.complex128), _device(b, torch.complex128)
    first = _noalloc(fk, a, ta, 0)
    _noalloc(fk, b, tb, 0)
    again = _noalloc(fk, a, ta, 0)
    assert np.array_equal(first[0], again[0]) and np.array_equal(first[1], again[1])


def _dummy_call(fk, fko, nb, nv, C, G, memory, **kw):
    """A call that the library must reject on the host: placeholder device
    pointers are never dereferenced."""
    p = fko.gpp_random(2, 1, 8, 2, seed=1)
    p.nb, p.nv, p.ncouls, p.ngpown = nb, nv, C, G
    p.inv_igp_index = (np.arange(G) * (C // G)).astype(np.int32)
    p.indinv = np.concatenate([np.arange(C), [C - 1]]).astype(np.int32)
    s, keep = _struct(fk, p, dict(aqsm=1, aqsn=1, I_eps=1, wtilde=1, vcoul=1))
    opt = fk.default_options()
    opt.memory = memory
    for k, v in kw.items():
        setattr(opt, k, v)
    out = ctypes.create_string_buffer(64)
    rc = fk.lib.gpp_kernel_noalloc(ctypes.byref(s), ctypes.byref(opt), ctypes.addressof(out),
                                   ctypes.addressof(out), None)
    del keep
    return rc


def test_noalloc_rejections_before_any_device_work(fk, fko):
    limit = int(fk.lib.gpp_kernel_noalloc_max_workspace())
    assert limit == 64 << 20
    # 512/2/32768/1638 FP64 fits; 1024 bands x 4096 columns does not
    assert fk.gpp_workspace_size(512, 2, 32768, 1638, 3) <= limit
    assert fk.gpp_workspace_size(1024, 2, 32768, 4096, 3) > limit
    assert _dummy_call(fk, fko, 1024, 2, 32768, 4096, fk.FK_MEM_DEVICE) == fk.FK_ERR_WORKSPACE
    assert "scratch" in fk.last_error()
    assert _dummy_call(fk, fko, 4, 1, 64, 8, fk.FK_MEM_HOST) == fk.FK_ERR_UNSUPPORTED
    assert _dummy_call(fk, fko, 4, 1, 64, 8, fk.FK_MEM_DEVICE, workspace=1, workspace_bytes=1 << 20) == fk.FK_ERR_ARG
    assert _dummy_call(fk, fko, 4, 5, 64, 8, fk.FK_MEM_DEVICE) == fk.FK_ERR_ARG   # nv > nb


def test_noalloc_rejects_acht_n1_loc(fk, fko):
    p = fko.gpp_random(2, 1, 8, 2, seed=1)
    s, keep = _struct(fk, p, dict(aqsm=1, aqsn=1, I_eps=1, wtilde=1, vcoul=1))
    opt = fk.default_options()
    opt.memory = fk.FK_MEM_DEVICE
    buf = ctypes.create_string_buffer(256)
    rc = fk.lib.gpp_kernel_noalloc(ctypes.byref(s), ctypes.byref(opt), ctypes.addressof(buf),
                                   ctypes.addressof(buf), ctypes.addressof(buf))
    assert rc == fk.FK_ERR_UNSUPPORTED and "acht_n1_loc" in fk.last_error()
    del keep
