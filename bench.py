#!/usr/bin/env python3
"""Benchmark of the GPP band sum on one MI355X.

A step is one gpp_kernel call (FP64, device memory, inputs resident in HBM)
over the 512/2/32768/20 workload: number_bands 512, nvband 2, ncouls 32768,
ngpown 1638, 3 frequencies, seeded synthetic inputs (aqs*/I_eps ~ U(-0.5,0.5)
complex, wtilde re ~ U(0.5,2), im ~ U(-0.01,0.01), vcoul ~ U(0.1,10)).

Prints ONE JSON line on stdout.  value = updates/s (n1 x igp x ig x iw per
second).  roofline = the fused GEMM+epilogue kernel's algorithmic FP64 flops
over its HIP-event-timed duration against the 78.6 TFLOP/s FP64 MFMA peak.
cpu_baseline = the reference's own gppKer program (oracle/_ref, built from
the reference sources) on a bounded band sample of the same shape, timed on
this host's cores; the CPU oracle restatement is used if it is absent.

More objects ride on the same line: "mixed" (the complex64-storage GPP
form on the same shape), "ff" (the FF frequency contraction at
number_bands 256, ncouls = ngpown = 4096, nFreq 32, FP64 MFMA, with its own
roofline against the FP64 MFMA peak) and "batched" (64 uneven GPP problems
in one graph launch: FP64 on the persistent tile kernel, tile configuration
chosen by the library's autotuner; mixed on the int8-split contraction).
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GPP updates/s (n1·igp·ig·iw) + % FP64/FP32 vector peak, FP64 vs mixed"
PEAK_FP64 = 78.6      # TFLOP/s, MI355X FP64 vector = FP64 MFMA (dense)
PEAK_FP32 = 157.3     # TFLOP/s, FP32 vector = FP32 MFMA
PEAK_INT8 = 5000.0    # TOP/s, int8 MFMA (dense)
PEAK_BF16 = 2500.0    # TFLOP/s, BF16 MFMA (dense)
NB, NV, NCOULS, NPG, NW = 512, 2, 32768, 20, 3
STREAMS = {"aqsm": 1, "aqsn": 2, "I_eps": 3, "wtilde": 4, "vcoul": 5}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def gpp_alg_flops(nb, nv, C, G, nw):
    """FP64 flops of the factorised evaluation (one reciprocal = 1 flop):
    8 nb G C for B = sum_n1 c[n1] aqsn[n1] (complex MAC), and per (igp, ig):
    8 nv (occupied-band contraction) + 9 (wtilde^2, Omega2) + nw x 42
    (pole factor 22, SX factor 4, two complex MACs 16)."""
    return 8.0 * nb * G * C + G * C * (8.0 * nv + 9.0 + 42.0 * nw)


def gpp_executed_flops(nb, nv, C, G, nw):
    """FP64 flops the kernel executes for the same work: the band product
    runs as three real products per complex multiply-add (the 3M method),
    6 nb G C, instead of the 8 nb G C that gpp_alg_flops counts."""
    return 6.0 * nb * G * C + G * C * (8.0 * nv + 9.0 + 42.0 * nw)


def gpp_reference_flops(nb, nv, C, G, nw):
    """The per-update count of the unfactorised loop with band-invariant work
    hoisted: 8U + 6 nb G C + 22 G C nw + 10 nv G C nw."""
    U = nb * G * C * nw
    return 8.0 * U + 6.0 * nb * G * C + 22.0 * G * C * nw + 10.0 * nv * G * C * nw


def import_fk():
    import importlib.util
    pkg = os.path.join(ROOT, "fortkernels-_amd")
    spec = importlib.util.spec_from_file_location("fortkernels_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["fortkernels_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def make_inputs(fk, torch, nb, C, G, seed, cdtype, wtilde_re):
    dev = torch.device("cuda:0")
    t = {}
    for name, rows, re_rng, im_rng in (("aqsm", nb, (-0.5, 0.5), (-0.5, 0.5)), ("aqsn", nb, (-0.5, 0.5), (-0.5, 0.5)),
                                       ("I_eps", G, (-0.5, 0.5), (-0.5, 0.5)),
                                       ("wtilde", G, wtilde_re, (-0.01, 0.01))):
        t[name] = torch.empty((rows, C), dtype=cdtype, device=dev)
        fk.fill_complex(t[name], seed, STREAMS[name], re_rng, im_rng)
    t["vcoul"] = torch.empty(C, dtype=torch.float64, device=dev)
    fk.fill_real(t["vcoul"], seed, STREAMS["vcoul"], 0.1, 10.0)
    torch.cuda.synchronize()
    return t


def index_maps(C, G):
    inv = (np.arange(G, dtype=np.int64) * C // G).astype(np.int32)     # i*ncouls/ngpown
    indinv = np.arange(C + 1, dtype=np.int32)
    indinv[C] = C - 1
    return inv, indinv


def wx_array(nw):
    return np.array([10.0 - 6.0 + 1.0 * ((iw + 1) - 2) for iw in range(nw)])


def run_gpp(fk, torch, precision, steps, warmup, seed=1):
    C, G = NCOULS, NCOULS // NPG
    cdt = torch.complex128 if precision == "fp64" else torch.complex64
    wre = (0.5, 2.0) if precision == "fp64" else (3.0, 5.0)
    log(f"{precision}: generating inputs nb={NB} nv={NV} ncouls={C} ngpown={G}")
    t = make_inputs(fk, torch, NB, C, G, seed, cdt, wre)
    inv, indinv = index_maps(C, G)
    wx = wx_array(NW)
    nbytes = fk.gpp_workspace_size(NB, NV, C, G, NW, precision=precision)
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0")
    stream = torch.cuda.current_stream()

    def step(time_kernel=False):
        return fk.gpp_kernel(NB, NV, C, G, wx, t["aqsm"], t["aqsn"], t["I_eps"], t["wtilde"], t["vcoul"],
                             inv, indinv, precision=precision, stream=stream, workspace=ws,
                             time_kernel=time_kernel)

    log(f"{precision}: warmup {warmup}")
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    log(f"{precision}: timing {steps} steps")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        out = step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = e0.elapsed_time(e1) / steps
    ach = out[0].cpu().numpy()
    asx = out[1].cpu().numpy()
    # dominant kernel: per-launch duration from HIP events on the launch stream
    kms = [step(time_kernel=True)[2] for _ in range(max(steps, 5))]
    torch.cuda.synchronize()
    res = dict(ms_per_step=1e3 * wall / steps, event_ms_per_step=ev_ms, kernel_ms=float(np.mean(kms)),
               kernel_ms_min=float(np.min(kms)), achtemp=ach, asxtemp=asx, C=C, G=G)
    if precision == "mixed":
        # the error of the mixed form against FP64 on the same complex64
        # values: the FP64 kernel (held to 1e-12 against the oracle by the
        # tests) on the inputs widened to complex128
        del ws
        w = {k: (v.to(torch.complex128) if v.dtype == torch.complex64 else v) for k, v in t.items()}
        del t
        ref = fk.gpp_kernel(NB, NV, C, G, wx, w["aqsm"], w["aqsn"], w["I_eps"], w["wtilde"], w["vcoul"], inv, indinv)
        torch.cuda.synchronize()
        ra, rs = ref[0].cpu().numpy(), ref[1].cpu().numpy()
        res["error_vs_fp64"] = {"achtemp": float(np.max(np.abs(ach - ra)) / np.max(np.abs(ra))),
                                "asxtemp": float(np.max(np.abs(asx - rs)) / np.max(np.abs(rs))),
                                "norm": "max|diff|/max|ref| over the frequencies, FP64 on the same complex64 values"}
        del w, ref
        torch.cuda.empty_cache()
    return res


FF_NB, FF_C, FF_G, FF_NF = 256, 4096, 4096, 32
FF_NV, FF_NE = 8, 10


def run_ff(fk, torch, steps, warmup, seed=4):
    """ff_kernel on I_epsR at 256/4096/4096/32, inputs resident in HBM."""
    dev = torch.device("cuda:0")
    nb, C, G, nf = FF_NB, FF_C, FF_G, FF_NF
    log(f"ff: generating inputs nb={nb} ncouls={C} ngpown={G} nfreq={nf} (I_epsR {16 * nf * G * C / 1e9:.2f} GB)")
    aqsm = torch.empty(nb, C, dtype=torch.complex128, device=dev)
    aqsn = torch.empty(nb, C, dtype=torch.complex128, device=dev)
    I = torch.empty(nf, G, C, dtype=torch.complex128, device=dev)
    v = torch.empty(C, dtype=torch.float64, device=dev)
    fk.fill_complex(aqsm, seed, 1, (-0.5, 0.5), (-0.5, 0.5))
    fk.fill_complex(aqsn, seed, 2, (-0.5, 0.5), (-0.5, 0.5))
    fk.fill_complex(I, seed, 6, (-0.5, 0.5), (-0.5, 0.5))
    fk.fill_real(v, seed, 5, 0.1, 10.0)
    inv = np.arange(G, dtype=np.int32)
    indinv = np.arange(C + 1, dtype=np.int32)
    indinv[C] = C - 1
    nbytes = fk.ff_workspace_size(nb, C, G, nf)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()

    def step(time_kernel=False):
        return fk.ff_kernel(nb, C, G, nf, aqsm, aqsn, I, v, inv, indinv, stream=stream, workspace=ws,
                            time_kernel=time_kernel)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    log(f"ff: timing {steps} steps")
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = [step(time_kernel=True)[3] for _ in range(max(steps, 3))]
    flops = 8.0 * G * nf * nb * C
    kernel_ms = float(np.mean(kms))
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    ms = 1e3 * wall / steps
    log(f"ff: {ms:.3f} ms/step, kernel {kernel_ms:.3f} ms, {achieved:.1f} TFLOP/s "
        f"({100 * achieved / PEAK_FP64:.1f}% of FP64 peak)")
    ref = step()
    out = {"workload": f"FF contraction {nb}/{C}/{G}/{nf} (number_bands/ncouls/ngpown/nFreq), FP64, I_epsR",
           "value": nb * G * C * nf / (ms * 1e-3), "unit": "complex MACs/s", "ms_per_step": ms,
           "roofline": {"bound": "mfma", "kernel": "ff_fused_kernel", "achieved": achieved, "peak": PEAK_FP64,
                        "unit": "TFLOP/s", "frac": achieved / PEAK_FP64,
                        "traffic": load_traffic("ff_fused_fp64_traffic.json"),
                        "alg_flops_per_launch": flops, "kernel_ms": kernel_ms,
                        "executed_flops_per_launch": 0.75 * flops,
                        "executed_frac": 0.75 * flops / (kernel_ms * 1e-3) / 1e12 / PEAK_FP64,
                        "note": "alg counts a complex multiply-add as 8 flops; the kernel runs 3 real MFMA "
                                "products per complex product (the 3M method), 3/4 of that"}}

    # the same contraction recomputed in split BF16 (three slices, six slice
    # products on MFMA, FP32 accumulation per ig chunk, the chunk sums
    # combined in FP64: FP32-class accuracy)
    nbytes = fk.ff_workspace_size(nb, C, G, nf, precision="mixed")
    ws_mix = torch.empty(nbytes, dtype=torch.uint8, device=dev)

    def step_mixed(time_kernel=False):
        return fk.ff_kernel(nb, C, G, nf, aqsm, aqsn, I, v, inv, indinv, stream=stream, workspace=ws_mix,
                            time_kernel=time_kernel, precision="mixed")

    for _ in range(warmup):
        step_mixed()
    torch.cuda.synchronize()
    log(f"ff split BF16: timing {steps} steps")
    t0 = time.perf_counter()
    for _ in range(steps):
        step_mixed()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = [step_mixed(time_kernel=True)[3] for _ in range(max(steps, 3))]
    mix = step_mixed()
    torch.cuda.synchronize()
    rel = lambda a, b: float((a - b).abs().max() / b.abs().max())  # noqa: E731
    err_t, err_v = rel(mix[0], ref[0]), rel(mix[1], ref[1])
    err_a = float(abs(complex(mix[2].cpu()[0]) - complex(ref[2].cpu()[0])) / abs(complex(ref[2].cpu()[0])))
    mms = 1e3 * wall / steps
    mkernel = float(np.mean(kms))
    bf16_flops = 6.0 * flops            # six slice products, each a full complex GEMM
    exec_flops = 0.75 * bf16_flops      # the kernel runs each as three real GEMMs (3M)
    mach = bf16_flops / (mkernel * 1e-3) / 1e12
    log(f"ff split BF16: {mms:.3f} ms/step ({ms / mms:.2f}x FP64), kernel {mkernel:.3f} ms, "
        f"{mach:.1f} TFLOP/s BF16 ({100 * mach / PEAK_BF16:.1f}% of BF16 MFMA peak), "
        f"error vs FP64 schDt {err_t:.2e} schDv {err_v:.2e} achsDtemp {err_a:.2e}")
    out["split_bf16"] = {
        "ms_per_step": mms, "speedup_vs_fp64": ms / mms, "value": nb * G * C * nf / (mms * 1e-3),
        "unit": "complex MACs/s",
        "error_vs_fp64": {"schDt": err_t, "schDv": err_v, "achsDtemp": err_a, "norm": "max|diff|/max|ref|"},
        "roofline": {"bound": "mfma", "kernel": "ff_split_gemm_kernel", "achieved": mach, "peak": PEAK_BF16,
                     "unit": "TFLOP/s", "frac": mach / PEAK_BF16,
                     "traffic": load_traffic("ff_split_gemm_traffic.json"),
                     "alg_flops_per_launch": bf16_flops, "kernel_ms": mkernel,
                     "executed_flops_per_launch": exec_flops,
                     "executed_frac": exec_flops / (mkernel * 1e-3) / 1e12 / PEAK_BF16,
                     "fp64_equivalent_tflops": flops / (mkernel * 1e-3) / 1e12,
                     "note": "alg counts each slice product as a complex GEMM of 8 flops per complex MAC; "
                             "the kernel runs it as three real products (the 3M method), 3/4 of that"}}
    del ws, ws_mix
    out["full_evaluation"] = run_ff_full(fk, torch, steps, warmup, aqsm, aqsn, I, v, inv, indinv, seed)
    return out


def run_ff_full(fk, torch, steps, warmup, aqsm, aqsn, IR, v, inv, indinv, seed):
    """ff_kernel_full: the whole FF evaluation of the Fortran program (static
    CH, SX, sigma 1-4, delta terms) at nb=256, nvband=8, 4096/4096/32,
    nfreqeval=10 with the program's axes; FP64 and split BF16."""
    dev = torch.device("cuda:0")
    nb, C, G, nf = FF_NB, FF_C, FF_G, FF_NF
    nv, ne = FF_NV, FF_NE
    IA = torch.empty(nf, G, C, dtype=torch.complex128, device=dev)
    fk.fill_complex(IA, seed, 7, (-0.5, 0.5), (-0.5, 0.5))
    ekq = -10.0 + np.arange(nb, dtype=np.float64)           # ffkernel.deslippe.best.f90:124-131
    grid = 2.0 * np.arange(nf, dtype=np.float64)
    pref = np.empty(nf)
    pref[:-1] = (grid[1:] - grid[:-1]) / 3.14
    pref[-1] = pref[-2]
    pref[0] *= 0.5
    pref[-1] *= 0.5
    brd = np.full(nf, 0.1j)
    stream = torch.cuda.current_stream()
    res, outs = {}, {}
    for prec in ("fp64", "mixed"):
        n = fk.ff_full_workspace_size(nb, nv, C, G, nf, ne, ekq, grid, precision=prec)
        ws = torch.empty(n, dtype=torch.uint8, device=dev)

        def step(time_kernel=False):
            return fk.ff_kernel_full(nb, nv, C, G, nf, ne, aqsm, aqsn, IR, IA, v, inv, indinv, ekq, grid, pref,
                                     brd, stream=stream, workspace=ws, time_kernel=time_kernel, precision=prec,
                                     want_schDt=False)

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        log(f"ff full evaluation {prec}: timing {steps} steps")
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kms = [step(time_kernel=True)["kernel_ms"] for _ in range(3)]
        outs[prec] = step()
        torch.cuda.synchronize()
        res[prec] = {"ms_per_step": 1e3 * wall / steps, "device_ms": float(np.mean(kms))}
        del ws
    rel = lambda a, b: float((a - b).abs().max() / b.abs().max())  # noqa: E731
    err = {k: rel(outs["mixed"][k], outs["fp64"][k]) for k in fk.FF_FULL_OUTPUTS + ("achsDtemp",)}
    res["split_bf16_error_vs_fp64"] = dict(err, norm="max|diff|/max|ref| over the evaluation frequencies")
    res["speedup_split_vs_fp64"] = res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"]
    res["workload"] = (f"FF evaluation {nb}/{nv}/{C}/{G}/{nf}/{ne} (number_bands/nvband/ncouls/ngpown/nFreq/"
                       "nfreqeval), I_epsR contracted in full, I_epsA over the bands and frequencies read")
    log(f"ff full evaluation: FP64 {res['fp64']['ms_per_step']:.3f} ms/step, split BF16 "
        f"{res['mixed']['ms_per_step']:.3f} ms/step; split error vs FP64 "
        + " ".join(f"{k} {v:.1e}" for k, v in err.items()))
    del IA
    return res


def batched_shapes(count=64, seed=3):
    """number_bands ~U{128..1024}, ncouls ~U{4096..32768} rounded to 64,
    ngpown = ncouls/20, nvband ~U{2..16}, frequencies ~U{3..11}."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nb = int(rng.integers(128, 1025))
        C = int(rng.integers(4096, 32769)) // 64 * 64
        out.append((nb, int(rng.integers(2, 17)), C, C // 20, int(rng.integers(3, 12))))
    return out


def run_batched(fk, torch, steps, singles=False):
    """64 problems as one batch, FP64 then mixed (complex64 storage, the
    int8-split contraction); inputs resident in HBM."""
    shapes = batched_shapes()
    updates = float(sum(nb * G * C * nw for nb, nv, C, G, nw in shapes))
    dev = torch.device("cuda:0")
    result = {"workload": "64 GPP problems: number_bands U{128..1024}, ncouls U{4096..32768}/64*64, "
                          "ngpown ncouls/20, nvband U{2..16}, frequencies U{3..11}, one graph launch",
              "updates": updates}
    keep = {}
    for prec, cdt in (("fp64", torch.complex128), ("mixed", torch.complex64)):
        log(f"batched {prec}: generating 64 problems ({updates:.3e} updates)")
        probs = []
        for k, (nb, nv, C, G, nw) in enumerate(shapes):
            seed = 100 + k
            t = {}
            for name, rows, re in (("aqsm", nb, (-0.5, 0.5)), ("aqsn", nb, (-0.5, 0.5)), ("I_eps", G, (-0.5, 0.5)),
                                   ("wtilde", G, (0.5, 2.0))):
                t[name] = torch.empty((rows, C), dtype=cdt, device=dev)
                fk.fill_complex(t[name], seed, STREAMS[name], re, (-0.01, 0.01) if name == "wtilde" else (-0.5, 0.5))
            t["vcoul"] = torch.empty(C, dtype=torch.float64, device=dev)
            fk.fill_real(t["vcoul"], seed, STREAMS["vcoul"], 0.1, 10.0)
            inv, indinv = index_maps(C, G)
            probs.append(dict(nb=nb, nv=nv, ncouls=C, ngpown=G, wx=wx_array(nw), inv_igp_index=inv, indinv=indinv, **t))
        torch.cuda.synchronize()
        log(f"batched {prec}: create (autotuning)")
        t0 = time.perf_counter()
        b = fk.GppBatch(probs, precision=prec, config=0)
        create_s = time.perf_counter() - t0
        b.run()
        torch.cuda.synchronize()
        log(f"batched {prec}: config {b.config}, timing {steps} runs")
        t0 = time.perf_counter()
        for _ in range(steps):
            ach, asx, _ = b.run()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        keep[prec] = [a.cpu().numpy() for a in ach]
        result[prec] = {"value": updates / (ms * 1e-3), "unit": "updates/s", "ms_per_step": ms, "config": b.config,
                        "create_s": create_s, "workspace_bytes": b.workspace_bytes}
        if prec == "mixed":
            # error against FP64 on the same complex64 values, on 8 problems
            worst = 0.0
            for k in range(0, len(probs), 8):
                p = probs[k]
                up = {n: p[n].to(torch.complex128) for n in ("aqsm", "aqsn", "I_eps", "wtilde")}
                ra, rs, _ = fk.gpp_kernel(p["nb"], p["nv"], p["ncouls"], p["ngpown"], p["wx"], up["aqsm"], up["aqsn"],
                                          up["I_eps"], up["wtilde"], p["vcoul"], p["inv_igp_index"], p["indinv"])
                for got, ref in ((ach[k], ra), (asx[k], rs)):
                    worst = max(worst, float((got - ref).abs().max() / ref.abs().max()))
                del up
            result[prec]["error_vs_fp64_same_values"] = {"max": worst, "problems": len(range(0, len(probs), 8)),
                                                         "norm": "max|diff|/max|ref| per output"}
        if singles:
            # the same 64 problems as 64 gpp_kernel calls back to back on one
            # stream (each call's own launches and workspace), one
            # synchronisation at the end; best of 2
            kw = {"precision": "mixed", "config": 0} if prec == "mixed" else {}
            best = None
            for _ in range(2):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for p in probs:
                    fk.gpp_kernel(p["nb"], p["nv"], p["ncouls"], p["ngpown"], p["wx"], p["aqsm"], p["aqsn"],
                                  p["I_eps"], p["wtilde"], p["vcoul"], p["inv_igp_index"], p["indinv"], **kw)
                torch.cuda.synchronize()
                t = 1e3 * (time.perf_counter() - t0)
                best = t if best is None else min(best, t)
            result[prec]["singles_back_to_back_ms"] = best
            result[prec]["batch_speedup_vs_singles"] = best / ms
            log(f"batched {prec}: 64 single calls back to back {best:.2f} ms (batch {ms:.2f} ms)")
        if prec == "fp64" and singles:
            # each problem alone through gpp_kernel, best tile configuration
            # per problem, fused-kernel time (kernel_ms)
            tot = 0.0
            per_cfg = {c: 0.0 for c in (1, 2, 3, 4)}
            wins = {c: 0 for c in (1, 2, 3, 4)}
            for p in probs:
                t = {c: min(fk.gpp_kernel(p["nb"], p["nv"], p["ncouls"], p["ngpown"], p["wx"], p["aqsm"],
                                          p["aqsn"], p["I_eps"], p["wtilde"], p["vcoul"], p["inv_igp_index"],
                                          p["indinv"], config=c, time_kernel=True)[2] for _ in range(3))
                     for c in (1, 2, 3, 4)}
                c_best = min(t, key=t.get)
                tot += t[c_best]
                wins[c_best] += 1
                for c in t:
                    per_cfg[c] += t[c]
            result[prec]["sum_of_single_problem_kernel_ms"] = tot
            result[prec]["single_problem_kernel_ms_per_config"] = per_cfg
            result[prec]["problems_fastest_at_config"] = wins
            log(f"batched {prec}: the 64 problems one by one, best configuration each: {tot:.2f} ms of kernels "
                f"(per configuration {', '.join(f'{c}: {v:.2f}' for c, v in per_cfg.items())}; "
                f"fastest at {', '.join(f'{c}: {n}' for c, n in wins.items())} problems)")
        if prec == "fp64":
            fl = sum(gpp_alg_flops(nb, nv, C, G, nw) for nb, nv, C, G, nw in shapes)
            result[prec]["achieved_tflops"] = fl / (ms * 1e-3) / 1e12
            result[prec]["frac_fp64_peak"] = result[prec]["achieved_tflops"] / PEAK_FP64
        log(f"batched {prec}: {ms:.2f} ms per batch, {updates / (ms * 1e-3):.3e} updates/s")
        b.close()
        del probs
        torch.cuda.empty_cache()
    err = max(float(np.max(np.abs(m - f)) / np.max(np.abs(f))) for m, f in zip(keep["mixed"], keep["fp64"]))
    result["mixed"]["speedup_vs_fp64"] = result["fp64"]["ms_per_step"] / result["mixed"]["ms_per_step"]
    result["mixed"]["max_rel_err_vs_fp64_inputs_achtemp"] = err
    result["mixed"]["error_note"] = ("max_rel_err_vs_fp64_inputs_achtemp includes the rounding of the inputs to "
                                     "complex64; error_vs_fp64_same_values does not")
    return result


def ff_batched_shapes(count=24, seed=5):
    """number_bands ~U{32..256}, ncouls ~U{512..3072} rounded to 64, ngpown
    = ncouls/2, nFreq ~U{8..32}: many small FF contractions."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nb = int(rng.integers(32, 257))
        C = int(rng.integers(512, 3073)) // 64 * 64
        out.append((nb, C, C // 2, int(rng.integers(8, 33))))
    return out


def run_ff_batched(fk, torch, steps):
    """24 uneven FF contractions: one ff_batch graph launch against the same
    problems through ff_kernel one after another (same stream, inputs
    resident in HBM), FP64 and split BF16, configuration autotuned."""
    shapes = ff_batched_shapes()
    dev = torch.device("cuda:0")
    flops = float(sum(8.0 * G * nf * nb * C for nb, C, G, nf in shapes))
    res = {"workload": "24 FF contractions: number_bands U{32..256}, ncouls U{512..3072}/64*64, ngpown ncouls/2, "
                       "nFreq U{8..32}", "flops": flops}
    probs = []
    for k, (nb, C, G, nf) in enumerate(shapes):
        t = {"aqsm": torch.empty(nb, C, dtype=torch.complex128, device=dev),
             "aqsn": torch.empty(nb, C, dtype=torch.complex128, device=dev),
             "I_eps": torch.empty(nf, G, C, dtype=torch.complex128, device=dev),
             "vcoul": torch.empty(C, dtype=torch.float64, device=dev)}
        for i, name in enumerate(("aqsm", "aqsn", "I_eps")):
            fk.fill_complex(t[name], 200 + k, 1 + i, (-0.5, 0.5), (-0.5, 0.5))
        fk.fill_real(t["vcoul"], 200 + k, 5, 0.1, 10.0)
        inv, indinv = index_maps(C, G)
        probs.append(dict(nb=nb, ncouls=C, ngpown=G, nfreq=nf, inv_igp_index=inv, indinv=indinv, **t))
    torch.cuda.synchronize()
    for prec in ("fp64", "mixed"):
        b = fk.FfBatch(probs, precision=prec, config=0)
        b.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run()
        torch.cuda.synchronize()
        ms_b = 1e3 * (time.perf_counter() - t0) / steps

        def singles():
            for p in probs:
                fk.ff_kernel(p["nb"], p["ncouls"], p["ngpown"], p["nfreq"], p["aqsm"], p["aqsn"], p["I_eps"],
                             p["vcoul"], p["inv_igp_index"], p["indinv"], config=b.config, precision=prec)

        singles()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            singles()
        torch.cuda.synchronize()
        ms_s = 1e3 * (time.perf_counter() - t0) / steps
        res[prec] = {"ms_per_step": ms_b, "config": b.config, "ms_one_call_per_problem": ms_s,
                     "speedup_vs_one_call_per_problem": ms_s / ms_b,
                     "achieved_tflops": flops / (ms_b * 1e-3) / 1e12}
        log(f"ff batched {prec}: {ms_b:.3f} ms per batch (config {b.config}) against {ms_s:.3f} ms "
            f"one call per problem")
        b.close()
    del probs
    torch.cuda.empty_cache()
    return res


def _host_cpu():
    """Model name, ISA flags and core counts of this host."""
    model, flags = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags") and not flags:
                    flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}, flags


def _reference_binary(flags):
    """The reference's gppKer built for the highest vector ISA level this
    host runs (AVX-512, else AVX2 + FMA, else generic x86-64)."""
    ref = os.path.join(ROOT, "oracle", "_ref")
    for name, need in (("gppKer_x86v4", "avx512f"), ("gppKer_x86v3", "avx2"), ("gppKer", None)):
        exe = os.path.join(ref, name)
        if os.access(exe, os.X_OK) and (need is None or need in flags):
            return exe, name
    return None, None


def cpu_baseline(bands, threads, runs=3):
    """The reference's own OpenMP gppKer program (BGW-Kernels/gppKer.cpp,
    built from its sources into oracle/_ref) on this host's cores: the
    median of `runs` runs of the full 128/2/8192/20 problem and of a
    `bands`-band sample of 512/2/32768/20 (the metric's shape); the timed
    region is the program's own band loop (its chrono line)."""
    host, flags = _host_cpu()
    exe, name = _reference_binary(flags)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread", OMP_PLACES="cores",
               OMP_STACKSIZE="16M")
    cases = [("128/2/8192/20, all 128 bands", 128, NV, 8192, NPG),
             (f"{bands} of the 512 bands of 512/2/32768/20", bands, NV, NCOULS, NPG)]
    out = {}
    for label, nb, nv, C, npg in cases:
        G = C // npg
        U = nb * G * C * NW
        secs = []
        for k in range(runs):
            if exe:
                log(f"cpu baseline: {name} {nb} {nv} {C} {npg} on {threads} threads (run {k + 1}/{runs})")
                text = subprocess.run([exe, str(nb), str(nv), str(C), str(npg)], env=env, capture_output=True,
                                      text=True, timeout=600, check=True).stdout
                secs.append(float(re.search(r"Chrono Time Taken \*+= ([0-9.eE+-]+)", text).group(1)))
            else:
                from oracle import fko
                log(f"cpu baseline: oracle port {nb}/{nv}/{C}/{G} on {threads} threads (run {k + 1}/{runs})")
                p = fko.gpp_random(nb, nv, C, G, seed=1)
                t0 = time.perf_counter()
                fko.gpp(p, extended=False if threads == 1 else True, nthreads=threads)
                secs.append(time.perf_counter() - t0)
        med = float(np.median(secs))
        out[label] = {"updates": U, "seconds": secs, "median_s": med, "updates_per_s": U / med}
    sample = out[cases[1][0]]
    return {"value": sample["updates_per_s"], "unit": "updates/s", "cores": threads,
            "kind": "reference" if exe else "port",
            "sample": f"{cases[1][0]}, nvband {NV}, 3 frequencies: median of {runs} runs "
                      f"({sample['updates']:.3e} updates, {sample['median_s']:.2f} s)",
            "binary": name or "oracle port (fko.gpp)",
            "build": "g++ -O3 -std=c++11 -fopenmp -march=" + {"gppKer_x86v4": "x86-64-v4", "gppKer_x86v3": "x86-64-v3",
                                                               "gppKer": "x86-64"}.get(name, "n/a"),
            "omp": "OMP_PROC_BIND=spread OMP_PLACES=cores OMP_STACKSIZE=16M",
            "host": host, "configs": out}


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-ff", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-singles", action="store_true",
                    help="also time the batched problems as gpp_kernel calls back to back (both precisions) "
                         "and each FP64 problem alone in its best configuration")
    ap.add_argument("--cpu-bands", type=int, default=128)
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0))),
                    help="host threads for the CPU baseline (default: OMP_NUM_THREADS, the host's CPU share "
                         "for this GPU, else every core this process may run on)")
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB three times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the GPP band sum is benchmarked per device (a multi-GPU run "
              "would shard ngpown columns and all-reduce 2*nfreq values); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    fk = import_fk()
    assert torch.cuda.is_available(), "no ROCm device"
    C, G = NCOULS, NCOULS // NPG
    U = NB * G * C * NW

    r = run_gpp(fk, torch, "fp64", args.steps, args.warmup)
    f_alg = gpp_alg_flops(NB, NV, C, G, NW)
    f_exec = gpp_executed_flops(NB, NV, C, G, NW)
    achieved = f_alg / (r["kernel_ms"] * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": U / (r["ms_per_step"] * 1e-3),
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "GPP band sum 512/2/32768/20 (ngpown 1638, 3 frequencies), FP64",
                   "number_bands": NB, "nvband": NV, "ncouls": C, "ngpown": G, "nfreq": NW},
        "roofline": {"bound": "mfma", "kernel": "gpp_fused_kernel", "achieved": achieved, "peak": PEAK_FP64,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64,
                     "traffic": load_traffic("gpp_fused_fp64_traffic.json"),
                     "alg_flops_per_launch": f_alg, "kernel_ms": r["kernel_ms"],
                     "executed_flops_per_launch": f_exec,
                     "executed_tflops": f_exec / (r["kernel_ms"] * 1e-3) / 1e12,
                     "executed_frac": f_exec / (r["kernel_ms"] * 1e-3) / 1e12 / PEAK_FP64,
                     "note": "alg counts a complex multiply-add as 8 flops; the kernel runs the band product "
                             "as 3 real MFMA products (6 flops), so executed = the MFMA and VALU work done"},
        "reference_equivalent_tflops": gpp_reference_flops(NB, NV, C, G, NW) / (r["ms_per_step"] * 1e-3) / 1e12,
        "event_ms_per_step": r["event_ms_per_step"],
    }
    log(f"fp64: {line['ms_per_step']:.3f} ms/step, kernel {r['kernel_ms']:.3f} ms, "
        f"{achieved:.1f} TFLOP/s ({100 * achieved / PEAK_FP64:.1f}% of FP64 peak)")
    if not args.no_mixed:
        m = run_gpp(fk, torch, "mixed", max(5, args.steps // 2), args.warmup)
        # the int8 GEMM: 3 real products (Gauss) x 10 slice pairs of a real
        # GEMM ngpown x ncouls x number_bands (2 ops per multiply-add)
        i8_ops = 2.0 * 3 * 10 * NB * G * C
        i8 = i8_ops / (m["kernel_ms"] * 1e-3) / 1e12
        line["mixed"] = {"value": U / (m["ms_per_step"] * 1e-3), "unit": "updates/s", "ms_per_step": m["ms_per_step"],
                         "speedup_vs_fp64": r["ms_per_step"] / m["ms_per_step"],
                         "dtype": "complex64 inputs; int8 MFMA split band contraction, FP64 elsewhere",
                         "roofline": {"bound": "mfma", "kernel": "gpp_oz_gemm_kernel", "achieved": i8,
                                      "peak": PEAK_INT8, "unit": "TOP/s", "frac": i8 / PEAK_INT8,
                                      "traffic": load_traffic("gpp_oz_gemm_traffic.json"),
                                      "alg_ops_per_launch": i8_ops, "kernel_ms": m["kernel_ms"]},
                         "error_vs_fp64": m["error_vs_fp64"],
                         "workload": "same shape, complex64 storage, wtilde re ~ U(3,5)"}
        log(f"mixed: {m['ms_per_step']:.3f} ms/step ({line['mixed']['speedup_vs_fp64']:.2f}x vs fp64), error vs "
            f"FP64 achtemp {m['error_vs_fp64']['achtemp']:.2e} asxtemp {m['error_vs_fp64']['asxtemp']:.2e}")
    if not args.no_ff:
        line["ff"] = run_ff(fk, torch, max(3, args.steps // 4), 1)
    if not args.no_batched:
        line["batched"] = run_batched(fk, torch, 3, singles=args.batch_singles)
        line["ff_batched"] = run_ff_batched(fk, torch, 5)
    if args.calibrate_copy:
        n = args.calibrate_copy << 20
        src = torch.ones(n // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: 3 x {args.calibrate_copy} MiB")
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.cpu_bands, args.cpu_threads, args.cpu_runs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
