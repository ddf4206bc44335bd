"""The 64-problem GPP batch at its real sizes: number_bands U{128..1024},
ncouls U{4096..32768} rounded to 64, ngpown = ncouls/20, nvband U{2..16},
3..11 frequencies, a distinct seed per problem, the distributions of the
128/2/8192/20 workload (inputs generated on the device).

Each problem of the batch is checked against gpp_kernel on the same problem
with the same tile configuration (gpp_kernel is held to the oracle at 1e-12
by test_gpp_gpu.py); the per-problem semantics are those of the band loop
BGW-Kernels/gppKer.cpp:303-439.  The oracle itself checks the batch on the
two occupied bands of the largest and of the smallest problem at their full
ngpown x ncouls shapes.  Replays are bitwise identical."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _torch():
    import torch
    assert torch.cuda.is_available(), "no ROCm device visible"
    return torch


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def shapes64(seed=2026):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(64):
        nb = int(rng.integers(128, 1025))
        C = int(rng.integers(4096, 32769)) // 64 * 64
        out.append((nb, int(rng.integers(2, 17)), C, C // 20, int(rng.integers(3, 12))))
    return out


def _make(fk, fko, shape, seed, cdt):
    torch = _torch()
    nb, nv, C, G, nw = shape
    t = {}
    for name, rows, re in (("aqsm", nb, (-0.5, 0.5)), ("aqsn", nb, (-0.5, 0.5)), ("I_eps", G, (-0.5, 0.5)),
                           ("wtilde", G, (0.5, 2.0))):
        t[name] = torch.empty((rows, C), dtype=cdt, device="cuda:0")
        im = (-0.01, 0.01) if name == "wtilde" else (-0.5, 0.5)
        fk.fill_complex(t[name], seed, fko.STREAM[name], re, im)
    t["vcoul"] = torch.empty(C, dtype=torch.float64, device="cuda:0")
    fk.fill_real(t["vcoul"], seed, fko.STREAM["vcoul"], 0.1, 10.0)
    inv, indinv = fko.index_maps(C, G, "fortran")
    return dict(nb=nb, nv=nv, ncouls=C, ngpown=G, wx=fko.wx_array(nw), inv_igp_index=inv, indinv=indinv, **t)


@pytest.fixture(scope="module")
def problems64(fk, fko):
    torch = _torch()
    sh = shapes64()
    nbytes = sum(16 * (2 * nb * C + 2 * G * C) for nb, nv, C, G, nw in sh)
    print(f"64 problems: {sum(nb * G * C * nw for nb, nv, C, G, nw in sh):.3e} updates, "
          f"{nbytes / 2**30:.1f} GiB of FP64 inputs", flush=True)
    probs = [_make(fk, fko, s, 500 + k, torch.complex128) for k, s in enumerate(sh)]
    torch.cuda.synchronize()
    yield probs
    del probs
    torch.cuda.empty_cache()


def _single(fk, p, **kw):
    return fk.gpp_kernel(p["nb"], p["nv"], p["ncouls"], p["ngpown"], p["wx"], p["aqsm"], p["aqsn"], p["I_eps"],
                         p["wtilde"], p["vcoul"], p["inv_igp_index"], p["indinv"], **kw)


def test_batch64_fp64_equals_single_calls(fk, problems64):
    """Autotuned FP64 batch: every problem within 1e-14 of gpp_kernel with the
    same tile configuration (the same tiles and fixed-order sums; the batch
    evaluates 4 frequencies per epilogue pass), and two runs bitwise equal."""
    torch = _torch()
    b = fk.GppBatch(problems64, config=0)
    ach, asx, ms = b.run(time_it=True)
    ach = [a.cpu().numpy() for a in ach]
    asx = [s.cpu().numpy() for s in asx]
    print(f"FP64 batch of 64: config {b.config}, {ms:.1f} ms, workspace {b.workspace_bytes / 2**20:.0f} MiB",
          flush=True)
    worst, same = 0.0, 0
    for k, p in enumerate(problems64):
        a1, s1, _ = _single(fk, p, config=b.config)
        torch.cuda.synchronize()
        a1, s1 = a1.cpu().numpy(), s1.cpu().numpy()
        e = max(_rel(ach[k], a1), _rel(asx[k], s1))
        worst = max(worst, e)
        same += int(np.array_equal(ach[k], a1) and np.array_equal(asx[k], s1))
        assert e <= 1e-14, f"problem {k} {p['nb']}/{p['nv']}/{p['ncouls']}/{p['ngpown']} nw {len(p['wx'])}: {e:.2e}"
    print(f"FP64 batch vs gpp_kernel: worst {worst:.2e}, {same} of 64 bitwise equal", flush=True)
    a2, s2, _ = b.run()
    torch.cuda.synchronize()
    for k in range(64):
        assert np.array_equal(a2[k].cpu().numpy(), ach[k]) and np.array_equal(s2[k].cpu().numpy(), asx[k])
    b.close()


def test_batch64_two_band_slices_vs_oracle(fk, fko, problems64):
    """The oracle on the two occupied bands of the largest and the smallest
    problem, each at its full ngpown x ncouls shape, run as a batch."""
    torch = _torch()
    size = [p["nb"] * p["ngpown"] * p["ncouls"] for p in problems64]
    picks = [int(np.argmax(size)), int(np.argmin(size))]
    sl = []
    for k in picks:
        p = problems64[k]
        sl.append(dict(p, nb=2, nv=2, aqsm=p["aqsm"][:2].contiguous(), aqsn=p["aqsn"][:2].contiguous()))
    b = fk.GppBatch(sl, config=2)
    ach, asx, _ = b.run()
    torch.cuda.synchronize()
    for j, (k, p) in enumerate(zip(picks, sl)):
        o = fko.GppProblem(2, 2, p["ncouls"], p["ngpown"], p["wx"], p["aqsm"].cpu().numpy(),
                           p["aqsn"].cpu().numpy(), p["I_eps"].cpu().numpy(), p["wtilde"].cpu().numpy(),
                           p["vcoul"].cpu().numpy(), p["inv_igp_index"], p["indinv"])
        ra, rs, _ = fko.gpp(o, nthreads=0)
        ea, es = _rel(ach[j].cpu().numpy(), ra), _rel(asx[j].cpu().numpy(), rs)
        print(f"problem {k} ({p['ncouls']}/{p['ngpown']}, nw {len(p['wx'])}), 2 bands vs oracle: "
              f"{ea:.1e} {es:.1e}", flush=True)
        assert ea <= RTOL and es <= RTOL
    b.close()


@pytest.mark.parametrize("config", [0, 2])
def test_batch64_mixed(fk, fko, problems64, config):
    """Mixed batch of the same 64 problems stored as complex64 (config 0: the
    int8-split contraction; 2: the FP32 MFMA tile kernel): each problem
    against gpp_kernel's mixed result on the same configuration, and every
    fourth against FP64 gpp_kernel on the same complex64 values (widened)."""
    torch = _torch()
    lo = [dict(p, **{k: p[k].to(torch.complex64) for k in ("aqsm", "aqsn", "I_eps", "wtilde")})
          for p in problems64]
    b = fk.GppBatch(lo, precision="mixed", config=config)
    assert b.config == config
    ach, asx, ms = b.run(time_it=True)
    ach = [a.cpu().numpy() for a in ach]
    asx = [s.cpu().numpy() for s in asx]
    print(f"mixed batch of 64: config {b.config}, {ms:.1f} ms, workspace {b.workspace_bytes / 2**20:.0f} MiB",
          flush=True)
    worst_fp64, worst_single = 0.0, 0.0
    # every problem against its single call (each one's tiles are walked by
    # the batch's own schedule); every fourth also against FP64
    for k in range(64):
        p = lo[k]
        m1, ms1, _ = _single(fk, p, precision="mixed", config=b.config)
        torch.cuda.synchronize()
        m1, ms1 = m1.cpu().numpy(), ms1.cpu().numpy()
        worst_single = max(worst_single, _rel(ach[k], m1), _rel(asx[k], ms1))
        if k % 4 == 0:
            wide = dict(p, **{n: p[n].to(torch.complex128) for n in ("aqsm", "aqsn", "I_eps", "wtilde")})
            ra, rs, _ = _single(fk, wide)
            torch.cuda.synchronize()
            worst_fp64 = max(worst_fp64, _rel(ach[k], ra.cpu().numpy()), _rel(asx[k], rs.cpu().numpy()))
            del wide
    bar = 1e-8 if b.config == 0 else 1e-5
    print(f"mixed batch: worst vs FP64 on the same values {worst_fp64:.2e} (bar {bar:.0e}), "
          f"vs gpp_kernel mixed {worst_single:.2e}", flush=True)
    assert worst_fp64 <= bar
    assert worst_single <= 1e-14
    b.close()
    del lo
    torch.cuda.empty_cache()
