"""C-ABI library: it loads, exports every symbol the header declares, and
rejects bad problems on the host side (no GPU needed for any of these)."""
import ctypes
import re

import numpy as np
import pytest


def _header_symbols(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, re.M)
    return sorted(set(n for n in names if n not in ("if", "sizeof")))


def test_library_exports_every_declared_symbol(fk):
    declared = _header_symbols(fk.INCLUDE)
    assert "gpp_kernel" in declared and "ff_kernel" in declared
    for name in declared:
        assert hasattr(fk.lib, name), f"missing export {name}"
    assert sorted(fk.EXPORTS) == declared


def test_abi_version_and_target(fk):
    assert fk.lib.fk_abi_version() == 3
    assert fk.lib.fk_build_target() == b"gfx950"


def test_default_options(fk):
    o = fk.default_options()
    assert (o.memory, o.precision, o.mask, o.device, o.config) == (fk.FK_MEM_HOST, fk.FK_PREC_FP64, 1, -1, 0)
    assert not o.stream and not o.workspace and not o.kernel_ms


def _small(fko):
    return fko.gpp_random(4, 2, 64, 5, seed=1)


def test_gpp_rejects_bad_sizes_before_touching_a_device(fk, fko):
    p = _small(fko)
    with pytest.raises(fk.FkError) as e:
        fk.gpp_kernel(p.nb, p.nb + 1, p.ncouls, p.ngpown, p.wx, p.aqsm, p.aqsn, p.I_eps, p.wtilde, p.vcoul,
                      p.inv_igp_index, p.indinv)
    assert e.value.code == fk.FK_ERR_ARG
    with pytest.raises(fk.FkError) as e:
        fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, np.zeros(17), p.aqsm, p.aqsn, p.I_eps, p.wtilde, p.vcoul,
                      p.inv_igp_index, p.indinv)
    assert e.value.code == fk.FK_ERR_ARG


@pytest.mark.parametrize("bad", ["indigp_high", "indigp_neg", "igp_eq_ncouls"])
def test_gpp_rejects_out_of_range_column_maps(fk, fko, bad):
    """Index maps the reference would read out of bounds (BGW-Kernels/gppKer.cpp:322-328)."""
    p = _small(fko)
    inv, ind = p.inv_igp_index.copy(), p.indinv.copy()
    if bad == "indigp_high":
        inv[1] = p.ncouls + 1
    elif bad == "indigp_neg":
        inv[0] = -1
    else:
        ind[inv[2]] = p.ncouls
    with pytest.raises(fk.FkError) as e:
        fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, p.aqsm, p.aqsn, p.I_eps, p.wtilde, p.vcoul, inv, ind)
    assert e.value.code == fk.FK_ERR_INDEX
    assert "ncouls" in fk.last_error()


def test_gpp_workspace_size_grows_with_problem(fk):
    small = fk.gpp_workspace_size(128, 2, 8192, 409, 3)
    big = fk.gpp_workspace_size(512, 2, 32768, 1638, 3)
    assert 0 < small < big
    assert big < 64 << 20   # coefficient matrix + per-tile partials only


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpp_workspace_size_with_acht_n1(fk, precision):
    """gpp_kernel_ex_workspace_size adds the per-band output's scratch (its
    G x ncouls pole factors, the widened band arrays of complex64 inputs,
    the row contraction's own) only when acht_n1_loc is asked for."""
    nb, C, G = 512, 32768, 1638
    base = fk.gpp_workspace_size(nb, 2, C, G, 3, precision=precision)
    n1 = fk.gpp_workspace_size(nb, 2, C, G, 3, precision=precision, acht_n1=True)
    extra = 16 * G * C + (2 * 16 * nb * C if precision == "mixed" else 0)
    assert n1 >= base + extra
    assert n1 < base + extra + (64 << 20)


def test_complex_helpers_host(fk):
    rng = np.random.default_rng(0)
    a = rng.normal(size=64) + 1j * rng.normal(size=64)
    b = rng.normal(size=64) + 1j * rng.normal(size=64)
    s = rng.uniform(0.5, 2.0, 64)
    acc = rng.normal(size=64) + 1j * rng.normal(size=64)
    ap = lambda op, **kw: fk.complex_apply(op, **kw)  # noqa: E731
    assert np.allclose(ap("mul", a=a, b=b), a * b, rtol=1e-15)
    assert np.array_equal(ap("conj", a=a), np.conj(a))
    assert np.allclose(ap("abs2", a=a).real, np.abs(a) ** 2, rtol=1e-15)
    assert np.allclose(ap("abs", a=a).real, np.abs(a), rtol=1e-15)
    assert np.allclose(ap("scale", a=a, s=s), a * s, rtol=1e-15)
    assert np.allclose(ap("fma", a=a, b=b, out=acc), acc + a * b, rtol=1e-14)
    assert np.allclose(ap("fms", a=a, b=b, out=acc), acc - a * b, rtol=1e-14)
    assert np.allclose(ap("real_minus", a=a, s=s), s - a, rtol=1e-15)
    assert np.allclose(ap("real_plus", a=a, s=s), s + a, rtol=1e-15)
    assert np.allclose(ap("div", a=a, b=b), a / b, rtol=1e-14)
    assert np.allclose(ap("div_real", a=a, s=s), a / s, rtol=1e-15)
    assert np.allclose(ap("square", a=a), a * a, rtol=1e-15)
    assert np.allclose(ap("add", a=a, b=b), a + b, rtol=1e-15)
    assert np.allclose(ap("sub", a=a, b=b), a - b, rtol=1e-15)
    # identities: conj(ab) = conj(a) conj(b); a / a = 1; |a|^2 = a conj(a)
    assert np.allclose(ap("conj", a=ap("mul", a=a, b=b)), ap("mul", a=np.conj(a), b=np.conj(b)), rtol=1e-15)
    assert np.allclose(ap("div", a=a, b=a), np.ones(64), rtol=1e-15, atol=1e-15)
    assert np.allclose(ap("mul", a=a, b=np.conj(a)).real, ap("abs2", a=a).real, rtol=1e-15)


def test_ff_rejects_bad_problems_before_touching_a_device(fk, fko):
    """ngpown > ncouls is the reference program's usage error (FFKer_cori/ffKernel.cpp:320-325);
    achs_freq must name one of the frequencies; sizes must be positive."""
    p = fko.ff_random(4, 32, 8, 2, seed=1)
    args = (p.aqsm, p.aqsn, p.I_eps, p.vcoul, p.inv_igp_index, p.indinv)
    with pytest.raises(fk.FkError) as e:
        fk.ff_kernel(p.nb, 8, 32, p.nfreq, p.aqsm, p.aqsn, p.I_eps, p.vcoul, np.zeros(32, np.int32), p.indinv)
    assert e.value.code == fk.FK_ERR_ARG and "cannot be greater" in fk.last_error()
    for kw in ({"achs_freq": 2}, {"achs_freq": -1}):
        with pytest.raises(fk.FkError) as e:
            fk.ff_kernel(p.nb, p.ncouls, p.ngpown, p.nfreq, *args, **kw)
        assert e.value.code == fk.FK_ERR_ARG and "achs_freq" in fk.last_error()
    with pytest.raises(fk.FkError) as e:
        fk.ff_kernel(0, p.ncouls, p.ngpown, p.nfreq, *args)
    assert e.value.code == fk.FK_ERR_ARG
    bad = p.inv_igp_index.copy()
    bad[0] = p.ncouls + 1
    with pytest.raises(fk.FkError) as e:
        fk.ff_kernel(p.nb, p.ncouls, p.ngpown, p.nfreq, p.aqsm, p.aqsn, p.I_eps, p.vcoul, bad, p.indinv)
    assert e.value.code == fk.FK_ERR_INDEX


def test_ff_workspace_size(fk):
    small = fk.ff_workspace_size(6, 40, 5, 3)
    big = fk.ff_workspace_size(256, 4096, 4096, 32)
    assert 0 < small < big < 64 << 20   # cf + column weights + per-tile partials only


def test_ff_mixed_workspace_and_config_range(fk, fko):
    """The split-BF16 path adds aqsn's slices (6 bytes per real) and per-chunk
    partials; its configs are 0..3 (ig chunk of 1024, 256, 512, 1024)."""
    nb, C, G, nf = 256, 4096, 4096, 32
    mixed = fk.ff_workspace_size(nb, C, G, nf, precision="mixed")
    slices = (2 * C // 16) * (nb // 32) * 3 * 1024     # K'/16 k-steps x 32-row fragments x 3 slices x 1 KiB
    partials = nf * (C // 1024) * (G // 64) * nb * 2 * 16
    assert mixed >= fk.ff_workspace_size(nb, C, G, nf) - 1 and mixed >= slices + partials
    assert fk.ff_workspace_size(nb, C, G, nf, config=3, precision="mixed") == mixed
    assert fk.ff_workspace_size(nb, C, G, nf, config=1, precision="mixed") > mixed
    p = fko.ff_random(4, 32, 8, 2, seed=1)
    with pytest.raises(fk.FkError) as e:
        fk.ff_problem(p, precision="mixed", config=4)
    assert e.value.code == fk.FK_ERR_ARG and "config" in fk.last_error()


def test_batch_rejects_host_memory_and_bad_sizes(fk, fko):
    """Argument checks of gpp_batch_create run before any device work."""
    import ctypes
    p = fko.gpp_random(4, 2, 64, 5, seed=1)
    wx = np.ascontiguousarray(p.wx)
    s = fk.GppProblem(p.nb, p.nv, p.ncouls, p.ngpown, len(wx), wx.ctypes.data, 1, 1, 1, 1, 1,
                      p.inv_igp_index.ctypes.data, p.indinv.ctypes.data, 1e-6, 0.25, 2.5e5, 1.0)
    arr = (ctypes.POINTER(fk.GppProblem) * 1)(ctypes.pointer(s))
    h = ctypes.c_void_p()
    opt = fk.default_options()          # host memory
    assert fk.lib.gpp_batch_create(1, arr, ctypes.byref(opt), ctypes.byref(h)) == fk.FK_ERR_UNSUPPORTED
    assert "FK_MEM_DEVICE" in fk.last_error()
    opt.memory = fk.FK_MEM_DEVICE
    assert fk.lib.gpp_batch_create(0, arr, ctypes.byref(opt), ctypes.byref(h)) == fk.FK_ERR_ARG
    opt.config = 9
    assert fk.lib.gpp_batch_create(1, arr, ctypes.byref(opt), ctypes.byref(h)) == fk.FK_ERR_ARG
    assert not h.value


def test_ff_workspace_grows_with_ig_split(fk):
    """Few FP64 tiles and a long ig range split the range into parts with a
    partial each: 70 bands x 40 columns x 2 frequencies at ncouls 3072 takes
    8 parts (64x64 tiles: 2 x 1 x 2 tiles; 256 ring passes of 3 k-steps), so
    the partials grow 8x over ncouls 256, whose 22 passes are too few to
    split (a part keeps at least 16); nothing else in the workspace depends
    on ncouls."""
    nb, G, nf = 70, 40, 2
    part = 2 * 16 * nf * nb          # bytes of partials per column tile
    short = fk.ff_workspace_size(nb, 256, G, nf, config=1)
    long_ = fk.ff_workspace_size(nb, 3072, G, nf, config=1)
    assert abs((long_ - short) - 7 * part) <= 256        # pieces are 256-byte aligned


def test_ff_full_workspace_and_checks(fk, fko):
    """ff_kernel_full's workspace covers the I_epsR contraction and, for the
    bands and frequencies the SX / delta terms read, the I_epsA one; its
    argument checks run before any device work."""
    nb, C, G, nf, ne = 64, 512, 256, 16, 10
    ekq, grid, _, _ = fko.ff_program_axes(nb, nf)
    base = fk.ff_workspace_size(nb, C, G, nf)
    # no occupied band and every wx < 0: I_epsA is never read
    none = fk.ff_full_workspace_size(nb, 0, C, G, nf, ne, ekq + 100.0, grid)
    occ = fk.ff_full_workspace_size(nb, 8, C, G, nf, ne, ekq, grid)
    assert none >= base and occ > none
    assert fk.ff_full_workspace_size(nb, 8, C, G, 1, ne, ekq, grid[:1]) == 0          # nfreq >= 2
    bad = grid.copy()
    bad[3] = bad[2]
    assert fk.ff_full_workspace_size(nb, 8, C, G, nf, ne, ekq, bad) == 0               # strictly increasing
    assert fk.ff_full_workspace_size(nb, nb + 1, C, G, nf, ne, ekq, grid) == 0         # nv <= nb
    assert fk.ff_full_workspace_size(nb, 8, C, G, nf, 0, ekq, grid) == 0               # nfreqeval >= 1
    assert fk.ff_full_workspace_size(nb, 8, C, G, nf, ne, ekq, grid, precision="mixed") > 0
    p = fko.ff_full_random(6, 2, 16, 4, 3, 4, seed=1)
    p.dFreqGrid = p.dFreqGrid[::-1].copy()
    with pytest.raises(fk.FkError) as e:
        fk.ff_full_problem(p)
    assert e.value.code == fk.FK_ERR_ARG and "increasing" in fk.last_error()


def test_ff_full_semantics_checks(fk, fko):
    """The C++ port's semantics read I_epsA for other bands (wx' without iw,
    SX from I_epsR only), take freqevalmin / freqevalstep as int, and an
    unknown semantics value is rejected."""
    nb, C, G, nf, ne = 64, 512, 256, 16, 10
    ekq, grid, _, _ = fko.ff_program_axes(nb, nf)
    ws = {s: fk.ff_full_workspace_size(nb, 8, C, G, nf, ne, ekq, grid, semantics=s) for s in (0, 1, 2)}
    assert all(v > 0 for v in ws.values()) and ws[1] == ws[2]
    # every band unoccupied and wx' <= 0: the port reads no I_epsA at all
    assert (fk.ff_full_workspace_size(nb, 0, C, G, nf, ne, ekq + 100.0, grid, semantics=1)
            == fk.ff_full_workspace_size(nb, 0, C, G, nf, ne, ekq + 100.0, grid, semantics=0))
    assert fk.ff_full_workspace_size(nb, 8, C, G, nf, ne, ekq, grid, semantics=3) == 0
    assert fk.ff_full_workspace_size(nb, 8, C, G, nf, ne, ekq, grid, semantics=-1) == 0
    big = dict(freqevalmin=3e9)
    assert fk.ff_full_workspace_size(nb, 8, C, G, nf, ne, ekq, grid, semantics=0, **big) > 0
    assert fk.ff_full_workspace_size(nb, 8, C, G, nf, ne, ekq, grid, semantics=1, **big) == 0
    assert fk.ff_full_workspace_size(nb, 8, C, G, nf, ne, ekq, grid, semantics=1, freqevalstep=float("nan")) == 0
    # (nfreqeval-2) * int(freqevalstep) overflows int32
    assert fk.ff_full_workspace_size(nb, 8, C, G, nf, 40000, ekq, grid, semantics=2, freqevalstep=1e5) == 0
    p = fko.ff_full_random(6, 2, 16, 4, 3, 4, seed=1)
    with pytest.raises(fk.FkError) as e:
        fk.ff_full_problem(p, semantics=7)
    assert e.value.code == fk.FK_ERR_ARG and "semantics" in fk.last_error()


def test_ff_batch_rejects_host_memory_and_bad_config(fk, fko):
    """Argument checks of ff_batch_create run before any device work."""
    import ctypes
    p = fko.ff_random(4, 32, 8, 2, seed=1)
    s = fk.FfProblem(p.nb, p.ncouls, p.ngpown, p.nfreq, 1, 1, 1, 1, p.inv_igp_index.ctypes.data,
                     p.indinv.ctypes.data, 0)
    arr = (ctypes.POINTER(fk.FfProblem) * 1)(ctypes.pointer(s))
    h = ctypes.c_void_p()
    opt = fk.default_options()          # host memory
    assert fk.lib.ff_batch_create(1, arr, ctypes.byref(opt), ctypes.byref(h)) == fk.FK_ERR_UNSUPPORTED
    assert "FK_MEM_DEVICE" in fk.last_error()
    opt.memory = fk.FK_MEM_DEVICE
    assert fk.lib.ff_batch_create(0, arr, ctypes.byref(opt), ctypes.byref(h)) == fk.FK_ERR_ARG
    opt.config = 9
    assert fk.lib.ff_batch_create(1, arr, ctypes.byref(opt), ctypes.byref(h)) == fk.FK_ERR_ARG
    assert not h.value


def test_package_refuses_to_import_without_the_library(tmp_path):
    """No fallback: a copy of the package whose lib/ holds no built library
    raises ImportError on import instead of computing anything elsewhere."""
    import importlib.util
    import os
    import shutil
    src = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fortkernels-_amd",
                       "__init__.py")
    pkg = tmp_path / "fk_copy"
    pkg.mkdir()
    shutil.copy(src, pkg / "__init__.py")
    spec = importlib.util.spec_from_file_location("fk_copy_without_lib", str(pkg / "__init__.py"),
                                                  submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    with pytest.raises(ImportError) as e:
        spec.loader.exec_module(mod)
    assert "shared library not built" in str(e.value)


def test_gppsum_form_rejections_before_any_device_work(fk, fko):
    """FK_GPP_GPPSUM runs FP64 with the default mask and configuration and
    gives achtemp / asxtemp only; everything else is refused on the host."""
    import ctypes
    p = fko.gpp_random(4, 2, 64, 5, seed=1)
    kw = dict(semantics=fk.FK_GPP_GPPSUM)
    p64 = fko.GppProblem(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, *[getattr(p, k).astype(np.complex64) for k in
                                                                  ("aqsm", "aqsn", "I_eps", "wtilde")],
                         p.vcoul, p.inv_igp_index, p.indinv)
    with pytest.raises(fk.FkError) as e:
        fk.gpp_problem(p64, precision="mixed", **kw)
    assert e.value.code == fk.FK_ERR_UNSUPPORTED and "gppsum" in fk.last_error()
    for extra, code, word in ((dict(mask=False), fk.FK_ERR_UNSUPPORTED, "gppsum"),
                              (dict(config=2), fk.FK_ERR_UNSUPPORTED, "gppsum"),
                              (dict(acht_n1=True), fk.FK_ERR_UNSUPPORTED, "acht_n1_loc"),
                              (dict(achstemp=True), fk.FK_ERR_UNSUPPORTED, "achstemp"),
                              (dict(semantics=7), fk.FK_ERR_ARG, "semantics")):
        with pytest.raises(fk.FkError) as e:
            fk.gpp_problem(p, **dict(kw, **extra))
        assert e.value.code == code and word in fk.last_error(), extra
    wx = np.ascontiguousarray(p.wx)
    s = fk.GppProblem(p.nb, p.nv, p.ncouls, p.ngpown, len(wx), wx.ctypes.data, 1, 1, 1, 1, 1,
                      p.inv_igp_index.ctypes.data, p.indinv.ctypes.data, 1e-6, 0.25, 2.5e5, 1.0, fk.FK_GPP_GPPSUM)
    arr = (ctypes.POINTER(fk.GppProblem) * 1)(ctypes.pointer(s))
    h = ctypes.c_void_p()
    opt = fk.default_options()
    opt.memory = fk.FK_MEM_DEVICE
    assert fk.lib.gpp_batch_create(1, arr, ctypes.byref(opt), ctypes.byref(h)) == fk.FK_ERR_UNSUPPORTED
    assert "gppsum" in fk.last_error() and not h.value
    # the workspace of the form: the canonical layout with 64 x 64 tiles plus one int per column
    assert fk.lib.gpp_kernel_workspace_size(ctypes.byref(s), ctypes.byref(opt)) > 0


def test_python_wrappers_check_buffers_before_the_library(fk, fko):
    """Wrong dtype, too few elements, a non-contiguous array, a CPU tensor,
    numpy mixed with torch, or an unknown precision raise in Python before
    any pointer reaches the library."""
    import torch
    p = fko.gpp_random(4, 2, 64, 5, seed=1)
    args = lambda **o: [o.get(k, getattr(p, k)) for k in ("aqsm", "aqsn", "I_eps", "wtilde", "vcoul")]  # noqa: E731
    call = lambda *a, **kw: fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, *a, p.inv_igp_index,  # noqa: E731
                                          p.indinv, **kw)
    cases = [
        (TypeError, args(aqsm=p.aqsm.astype(np.complex64)), {}),            # complex64 with fp64
        (TypeError, args(), dict(precision="mixed")),                         # complex128 with mixed
        (ValueError, args(), dict(precision="fp32")),                         # unknown precision
        (ValueError, args(I_eps=p.I_eps[:3]), {}),                            # too few elements
        (ValueError, args(aqsn=np.asfortranarray(p.aqsn)), {}),               # not C-contiguous
        (TypeError, args(vcoul=p.vcoul.astype(np.float32)), {}),
        (TypeError, args(wtilde=torch.from_numpy(p.wtilde)), {}),             # torch mixed with numpy
        (TypeError, [torch.from_numpy(np.ascontiguousarray(a)) for a in args()], {}),   # CPU tensors
    ]
    for exc, a, kw in cases:
        with pytest.raises(exc):
            call(*a, **kw)
    with pytest.raises(ValueError):
        fk.gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, *args(), p.inv_igp_index[:2], p.indinv)
    q = fko.ff_random(4, 32, 8, 2, seed=1)
    with pytest.raises(ValueError):
        fk.ff_kernel(q.nb, q.ncouls, q.ngpown, q.nfreq + 1, q.aqsm, q.aqsn, q.I_eps, q.vcoul, q.inv_igp_index,
                     q.indinv)
    with pytest.raises(TypeError):
        fk.ff_kernel(q.nb, q.ncouls, q.ngpown, q.nfreq, q.aqsm, q.aqsn, q.I_eps.astype(np.complex64), q.vcoul,
                     q.inv_igp_index, q.indinv)
    with pytest.raises(TypeError):
        fk.GppBatch([dict(nb=p.nb, nv=p.nv, ncouls=p.ncouls, ngpown=p.ngpown, wx=p.wx, aqsm=p.aqsm, aqsn=p.aqsn,
                          I_eps=p.I_eps, wtilde=p.wtilde, vcoul=p.vcoul, inv_igp_index=p.inv_igp_index,
                          indinv=p.indinv)])
    with pytest.raises(TypeError):
        fk.fill_real(np.zeros(4), 1, 1, 0.0, 1.0)
