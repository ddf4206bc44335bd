"""GPU parity of ff_kernel with FK_PREC_MIXED: the contraction in split BF16
on MFMA (three slices, FP32 accumulation over one ig chunk, FP64 across
chunks and in the contraction with conj(aqsm) and vcoul), against the FP64
oracle on the same complex128 inputs.

What pins it:
  * inputs that BF16 holds exactly (multiples of 1/8 in [-1, 1]): one slice,
    exact products, FP32 sums below 2^24 units -> the split path must agree
    with the FP64 oracle to FP64 rounding (<= 1e-13).  A wrong sign, a
    transposed operand, a misplaced fragment or a dropped tile fails this;
  * inputs with 9 or 10 significant bits and ncouls <= 32: exactly two
    slices, all four of their products kept, sums exact -> again <= 1e-13.
    This pins the cross-slice products and the residual signs;
  * random inputs ~U(-0.5, 0.5): the error against the FP64 oracle is stated
    and bounded, MIXED_RTOL (normwise, max |diff| / max |ref|); measured on
    MI355X 1.7e-7 .. 7e-7 for schDt and schDv, up to 1.6e-6 for achsDtemp.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXACT_RTOL = 1e-13
MIXED_RTOL = 5e-6


def _torch():
    import torch
    assert torch.cuda.is_available(), "no ROCm device visible"
    return torch


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _run(fk, p, **kw):
    torch = _torch()
    d = torch.device("cuda:0")
    f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(d).to(dt)  # noqa: E731
    out = fk.ff_kernel(p.nb, p.ncouls, p.ngpown, p.nfreq, f(p.aqsm, torch.complex128), f(p.aqsn, torch.complex128),
                       f(p.I_eps, torch.complex128), f(p.vcoul, torch.float64), p.inv_igp_index, p.indinv,
                       precision="mixed", **kw)
    torch.cuda.synchronize()
    schDt, schDv, achs, ms = out
    return schDt.cpu().numpy(), (None if schDv is None else schDv.cpu().numpy()), complex(achs.cpu()[0]), ms


def _grid_values(rng, shape, denom, top):
    """complex values (m + i n) / denom with integers |m|, |n| <= top"""
    m = rng.integers(-top, top + 1, size=shape)
    n = rng.integers(-top, top + 1, size=shape)
    return (m + 1j * n) / denom


@pytest.mark.parametrize("shape", [(6, 40, 5, 3), (70, 300, 130, 4), (257, 1100, 65, 2), (1, 1, 1, 1)])
def test_exact_on_bf16_inputs(fk, fko, shape):
    nb, C, G, nf = shape
    print(f"ff split exact shape {shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    p = fko.ff_random(nb, C, G, nf, seed=sum(shape))
    p.aqsn = _grid_values(rng, (nb, C), 8.0, 8)
    p.I_eps = _grid_values(rng, (nf, G, C), 8.0, 8)
    rt, rv, ra = fko.ff(p, achs_freq=nf - 1)
    gt, gv, ga, _ = _run(fk, p, achs_freq=nf - 1)
    assert _rel(gt, rt) <= EXACT_RTOL, f"schDt rel {_rel(gt, rt):.3e}"
    assert _rel(gv, rv) <= EXACT_RTOL, f"schDv rel {_rel(gv, rv):.3e}"
    assert abs(ga - ra) <= EXACT_RTOL * abs(ra)


@pytest.mark.parametrize("shape,bits", [((100, 32, 32, 3), 9), ((33, 8, 8, 2), 10)])
def test_exact_on_two_slice_inputs(fk, fko, shape, bits):
    """Values m / 2^bits, |m| < 2^bits: two slices; 2*ncouls products of
    2*bits bits sum below 2^24 units, so FP32 holds every partial sum."""
    nb, C, G, nf = shape
    print(f"ff split two-slice shape {shape}", flush=True)
    assert 2 * C * 4 ** bits <= 2 ** 24
    rng = np.random.default_rng(7 + sum(shape))
    p = fko.ff_random(nb, C, G, nf, seed=sum(shape))
    p.aqsn = _grid_values(rng, (nb, C), 2.0 ** bits, 2 ** bits - 1)
    p.I_eps = _grid_values(rng, (nf, G, C), 2.0 ** bits, 2 ** bits - 1)
    rt, rv, _ = fko.ff(p)
    gt, gv, _, _ = _run(fk, p)
    assert _rel(gt, rt) <= EXACT_RTOL, f"schDt rel {_rel(gt, rt):.3e}"
    assert _rel(gv, rv) <= EXACT_RTOL, f"schDv rel {_rel(gv, rv):.3e}"


@pytest.mark.parametrize("shape", [(6, 40, 5, 3), (70, 300, 130, 4), (129, 1000, 200, 3), (256, 2048, 128, 2)])
def test_error_vs_fp64_oracle(fk, fko, shape):
    nb, C, G, nf = shape
    print(f"ff split random shape {shape}", flush=True)
    p = fko.ff_random(nb, C, G, nf, seed=sum(shape))
    rt, rv, ra = fko.ff(p)
    gt, gv, ga, _ = _run(fk, p)
    et, ev = _rel(gt, rt), _rel(gv, rv)
    print(f"  split BF16 vs FP64 oracle: schDt {et:.2e}  schDv {ev:.2e}  achs {abs(ga - ra) / abs(ra):.2e}")
    assert et <= MIXED_RTOL and ev <= MIXED_RTOL
    assert abs(ga - ra) <= MIXED_RTOL * abs(ra)


@pytest.mark.parametrize("config", [1, 2])
def test_chunk_configs(fk, fko, config):
    """ig chunks of 256 and 512 (the default is 1024): the chunk partials sum
    to the same result."""
    p = fko.ff_random(40, 1500, 70, 2, seed=config)
    rt, _, _ = fko.ff(p)
    gt, _, _, _ = _run(fk, p, config=config)
    assert _rel(gt, rt) <= MIXED_RTOL


def test_permuted_and_skipped_columns(fk, fko):
    rng = np.random.default_rng(5)
    p = fko.ff_random(20, 96, 12, 2, seed=11)
    p.aqsn = _grid_values(rng, (p.nb, p.ncouls), 8.0, 8)
    p.I_eps = _grid_values(rng, (p.nfreq, p.ngpown, p.ncouls), 8.0, 8)
    inv = p.inv_igp_index.copy()
    inv[[1, 3]] = inv[[3, 1]]
    inv[5] = p.ncouls                      # indigp == ncouls -> igp = ncouls - 1
    ind = p.indinv.copy()
    ind[inv[7]] = -1                       # skipped column
    p.inv_igp_index, p.indinv = inv, ind
    rt, rv, _ = fko.ff(p)
    gt, gv, _, _ = _run(fk, p)
    assert _rel(gt, rt) <= EXACT_RTOL and _rel(gv, rv) <= EXACT_RTOL


def test_deterministic_host_memory_and_workspace(fk, fko):
    torch = _torch()
    p = fko.ff_random(50, 700, 90, 3, seed=9)
    a = _run(fk, p)
    b = _run(fk, p)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # host-memory call gives the same bits
    h = fk.ff_problem(p, precision="mixed")
    assert np.array_equal(h[0], a[0])
    # caller workspace, own stream, kernel timer
    n = fk.ff_workspace_size(p.nb, p.ncouls, p.ngpown, p.nfreq, precision="mixed")
    assert n > 0
    ws = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        c = _run(fk, p, stream=s, workspace=ws, time_kernel=True)
    assert c[3] is not None and c[3] > 0
    assert np.array_equal(c[0], a[0])


def test_full_size_error_vs_fp64(fk):
    """256/4096/4096/32: the split result against the FP64 MFMA result on the
    device (itself held to 1e-12 against the oracle by test_ff_gpu.py)."""
    torch = _torch()
    NB, C, G, NF = 256, 4096, 4096, 32
    d = torch.device("cuda:0")
    aqsm = torch.empty(NB, C, dtype=torch.complex128, device=d)
    aqsn = torch.empty(NB, C, dtype=torch.complex128, device=d)
    I = torch.empty(NF, G, C, dtype=torch.complex128, device=d)
    v = torch.empty(C, dtype=torch.float64, device=d)
    fk.fill_complex(aqsm, 4, 1, (-0.5, 0.5), (-0.5, 0.5))
    fk.fill_complex(aqsn, 4, 2, (-0.5, 0.5), (-0.5, 0.5))
    fk.fill_complex(I, 4, 6, (-0.5, 0.5), (-0.5, 0.5))
    fk.fill_real(v, 4, 5, 0.1, 10.0)
    inv = np.arange(G, dtype=np.int32)
    indinv = np.arange(C + 1, dtype=np.int32)
    indinv[C] = C - 1
    print("ff split full-size inputs filled", flush=True)
    ref = fk.ff_kernel(NB, C, G, NF, aqsm, aqsn, I, v, inv, indinv, time_kernel=True)
    mix = fk.ff_kernel(NB, C, G, NF, aqsm, aqsn, I, v, inv, indinv, time_kernel=True, precision="mixed")
    torch.cuda.synchronize()
    et = _rel(mix[0].cpu().numpy(), ref[0].cpu().numpy())
    ev = _rel(mix[1].cpu().numpy(), ref[1].cpu().numpy())
    ea = abs(complex(mix[2].cpu()[0]) - complex(ref[2].cpu()[0])) / abs(complex(ref[2].cpu()[0]))
    print(f"  full size: FP64 {ref[3]:.3f} ms, split BF16 {mix[3]:.3f} ms; "
          f"error schDt {et:.2e} schDv {ev:.2e} achs {ea:.2e}")
    assert et <= MIXED_RTOL and ev <= MIXED_RTOL and ea <= MIXED_RTOL
    del I
    torch.cuda.empty_cache()
