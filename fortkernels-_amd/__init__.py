"""Python binding of the MI355X GPP / FF library (include/fortkernels.h).

The product is the C ABI in lib/libfortkernels_amd.so (HIP, gfx950); this
module is a thin ctypes layer over it.  Arrays are numpy arrays (host
memory: the library copies, runs and copies back) or torch tensors on a
ROCm device (device memory: the call is asynchronous on the given stream).
There is no fallback: if the shared library is missing, importing this
module raises.

The names mirror the reference's driver arguments (BGW-Kernels/gppKer.cpp
main: number_bands, nvband, ncouls, nodes_per_group) and the C ABI entry
points gpp_kernel / ff_kernel.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfortkernels_amd.so")
INCLUDE = os.path.join(os.path.dirname(_HERE), "include", "fortkernels.h")

FK_OK, FK_ERR_ARG, FK_ERR_INDEX, FK_ERR_UNSUPPORTED, FK_ERR_HIP, FK_ERR_NOMEM, FK_ERR_WORKSPACE = 0, -1, -2, -3, -4, -5, -6
FK_MEM_HOST, FK_MEM_DEVICE = 0, 1
FK_PREC_FP64, FK_PREC_MIXED = 0, 1
COP = {"mul": 0, "conj": 1, "abs2": 2, "abs": 3, "scale": 4, "fma": 5, "fms": 6, "real_minus": 7,
       "real_plus": 8, "div": 9, "div_real": 10, "square": 11, "add": 12, "sub": 13}

# every extern "C" symbol of include/fortkernels.h
EXPORTS = ["fk_default_options", "gpp_kernel", "gpp_kernel_workspace_size", "ff_kernel",
           "ff_kernel_workspace_size", "fk_last_error", "fk_abi_version", "fk_build_target",
           "fk_fill_complex", "fk_fill_real", "fk_complex_apply", "fk_mfma_probe",
           "gpp_batch_create", "gpp_batch_run", "gpp_batch_config", "gpp_batch_workspace_bytes",
           "gpp_batch_destroy", "gpp_kernel_batched", "ff_kernel_full", "ff_kernel_full_workspace_size",
           "gpp_kernel_ex", "fk_gpp_instance_write", "fk_ff_instance_write", "fk_instance_open",
           "fk_instance_kind", "fk_instance_complex_bytes", "fk_instance_gpp", "fk_instance_ff", "fk_instance_close",
           "fk_instance_list_read", "fk_instance_list_free", "ff_batch_create", "ff_batch_run", "ff_batch_config",
           "ff_batch_workspace_bytes", "ff_batch_destroy", "gpp_kernel_noalloc", "gpp_kernel_noalloc_max_workspace",
           "gpp_kernel_ex_workspace_size"]


class FkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [("memory", ctypes.c_int32), ("precision", ctypes.c_int32), ("mask", ctypes.c_int32),
                ("device", ctypes.c_int32), ("stream", ctypes.c_void_p), ("workspace", ctypes.c_void_p),
                ("workspace_bytes", ctypes.c_size_t), ("config", ctypes.c_int32),
                ("kernel_ms", ctypes.POINTER(ctypes.c_float))]


class GppProblem(ctypes.Structure):
    _fields_ = [("number_bands", ctypes.c_int32), ("nvband", ctypes.c_int32), ("ncouls", ctypes.c_int32),
                ("ngpown", ctypes.c_int32), ("nfreq", ctypes.c_int32), ("wx", ctypes.c_void_p),
                ("aqsmtemp", ctypes.c_void_p), ("aqsntemp", ctypes.c_void_p), ("I_eps_array", ctypes.c_void_p),
                ("wtilde_array", ctypes.c_void_p), ("vcoul", ctypes.c_void_p),
                ("inv_igp_index", ctypes.c_void_p), ("indinv", ctypes.c_void_p),
                ("to1", ctypes.c_double), ("limittwo", ctypes.c_double), ("limitone", ctypes.c_double),
                ("occ", ctypes.c_double), ("semantics", ctypes.c_int32)]


# fk_gpp_semantics
FK_GPP_CANONICAL, FK_GPP_GPPSUM = 0, 1


class FfProblem(ctypes.Structure):
    _fields_ = [("number_bands", ctypes.c_int32), ("ncouls", ctypes.c_int32), ("ngpown", ctypes.c_int32),
                ("nfreq", ctypes.c_int32), ("aqsmtemp", ctypes.c_void_p), ("aqsntemp", ctypes.c_void_p),
                ("I_eps_array", ctypes.c_void_p), ("vcoul", ctypes.c_void_p),
                ("inv_igp_index", ctypes.c_void_p), ("indinv", ctypes.c_void_p), ("achs_freq", ctypes.c_int32)]


class GppOutputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("achtemp", "asxtemp", "acht_n1_loc", "achstemp")]


class FfFullProblem(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_int32) for n in ("number_bands", "nvband", "ncouls", "ngpown", "nfreq", "nfreqeval")]
                + [(n, ctypes.c_void_p) for n in ("aqsmtemp", "aqsntemp", "I_epsR_array", "I_epsA_array", "vcoul",
                                                   "inv_igp_index", "indinv", "ekq", "dFreqGrid", "pref",
                                                   "dFreqBrd")]
                + [(n, ctypes.c_double) for n in ("pref_zb", "freqevalmin", "freqevalstep", "occ")]
                + [("achs_freq", ctypes.c_int32), ("semantics", ctypes.c_int32)])


# fk_ff_semantics
FK_FF_FORTRAN, FK_FF_CPP_PORT, FK_FF_CPP_PORT_ZERO_SCHDT = 0, 1, 2
FF_FULL_OUTPUTS = ("asxDtemp", "achDtemp", "achDtemp_cor", "achDtemp_corb", "ach2Dtemp")


class FfFullOutputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("achsDtemp",) + FF_FULL_OUTPUTS + ("schDt",)]


def _load():
    # torch (when installed) ships its own HIP runtime whose SONAME is the
    # libamdhip64.so.7 this library needs.  Loading torch first makes the
    # dynamic loader reuse that runtime for the library, so device pointers
    # and streams from torch are valid here; loading the library first would
    # put a second, separate HIP runtime into the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"fortkernels-_amd: shared library not built: {LIB_PATH} "
                          "(run `make -C fortkernels-_amd` or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    P, I32, I64, U64, D, SZ = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64,
                               ctypes.c_double, ctypes.c_size_t)
    L.fk_default_options.restype = Options
    L.fk_default_options.argtypes = []
    L.gpp_kernel.restype = ctypes.c_int
    L.gpp_kernel.argtypes = [ctypes.POINTER(GppProblem), ctypes.POINTER(Options), P, P, P]
    L.gpp_kernel_workspace_size.restype = SZ
    L.gpp_kernel_workspace_size.argtypes = [ctypes.POINTER(GppProblem), ctypes.POINTER(Options)]
    L.gpp_kernel_noalloc.restype = ctypes.c_int
    L.gpp_kernel_noalloc.argtypes = [ctypes.POINTER(GppProblem), ctypes.POINTER(Options), P, P, P]
    L.gpp_kernel_noalloc_max_workspace.restype = SZ
    L.gpp_kernel_noalloc_max_workspace.argtypes = []
    L.ff_kernel.restype = ctypes.c_int
    L.ff_kernel.argtypes = [ctypes.POINTER(FfProblem), ctypes.POINTER(Options), P, P, P]
    L.ff_kernel_workspace_size.restype = SZ
    L.ff_kernel_workspace_size.argtypes = [ctypes.POINTER(FfProblem), ctypes.POINTER(Options)]
    L.fk_gpp_instance_write.restype = ctypes.c_int
    L.fk_gpp_instance_write.argtypes = [ctypes.c_char_p, ctypes.POINTER(GppProblem), I32]
    L.fk_ff_instance_write.restype = ctypes.c_int
    L.fk_ff_instance_write.argtypes = [ctypes.c_char_p, ctypes.POINTER(FfFullProblem)]
    L.fk_instance_open.restype = ctypes.c_int
    L.fk_instance_open.argtypes = [ctypes.c_char_p, ctypes.POINTER(P)]
    L.fk_instance_kind.restype = ctypes.c_int
    L.fk_instance_kind.argtypes = [P]
    L.fk_instance_complex_bytes.restype = I32
    L.fk_instance_complex_bytes.argtypes = [P]
    L.fk_instance_gpp.restype = ctypes.POINTER(GppProblem)
    L.fk_instance_gpp.argtypes = [P]
    L.fk_instance_ff.restype = ctypes.POINTER(FfFullProblem)
    L.fk_instance_ff.argtypes = [P]
    L.fk_instance_close.restype = None
    L.fk_instance_close.argtypes = [P]
    L.fk_instance_list_read.restype = ctypes.c_int
    L.fk_instance_list_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_char_p)),
                                        ctypes.POINTER(I32)]
    L.fk_instance_list_free.restype = None
    L.fk_instance_list_free.argtypes = [ctypes.POINTER(ctypes.c_char_p), I32]
    L.ff_batch_create.restype = ctypes.c_int
    L.ff_batch_create.argtypes = [I32, ctypes.POINTER(ctypes.POINTER(FfProblem)), ctypes.POINTER(Options),
                                  ctypes.POINTER(P)]
    L.ff_batch_run.restype = ctypes.c_int
    L.ff_batch_run.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P), P,
                               ctypes.POINTER(ctypes.c_float)]
    L.ff_batch_config.restype = ctypes.c_int
    L.ff_batch_config.argtypes = [P]
    L.ff_batch_workspace_bytes.restype = SZ
    L.ff_batch_workspace_bytes.argtypes = [P]
    L.ff_batch_destroy.restype = None
    L.ff_batch_destroy.argtypes = [P]
    L.gpp_kernel_ex.restype = ctypes.c_int
    L.gpp_kernel_ex.argtypes = [ctypes.POINTER(GppProblem), ctypes.POINTER(Options), ctypes.POINTER(GppOutputs)]
    L.gpp_kernel_ex_workspace_size.restype = SZ
    L.gpp_kernel_ex_workspace_size.argtypes = [ctypes.POINTER(GppProblem), ctypes.POINTER(Options),
                                               ctypes.POINTER(GppOutputs)]
    L.ff_kernel_full.restype = ctypes.c_int
    L.ff_kernel_full.argtypes = [ctypes.POINTER(FfFullProblem), ctypes.POINTER(Options),
                                 ctypes.POINTER(FfFullOutputs)]
    L.ff_kernel_full_workspace_size.restype = SZ
    L.ff_kernel_full_workspace_size.argtypes = [ctypes.POINTER(FfFullProblem), ctypes.POINTER(Options)]
    L.fk_last_error.restype = ctypes.c_char_p
    L.fk_last_error.argtypes = []
    L.fk_abi_version.restype = ctypes.c_int
    L.fk_build_target.restype = ctypes.c_char_p
    L.fk_fill_complex.restype = ctypes.c_int
    L.fk_fill_complex.argtypes = [P, I64, U64, U64, D, D, D, D, I32, P]
    L.fk_fill_real.restype = ctypes.c_int
    L.fk_fill_real.argtypes = [P, I64, U64, U64, D, D, P]
    L.fk_complex_apply.restype = ctypes.c_int
    L.fk_complex_apply.argtypes = [I32, I64, P, P, P, P, I32]
    L.fk_mfma_probe.restype = ctypes.c_int
    L.fk_mfma_probe.argtypes = [I32, P]
    PP = ctypes.POINTER(ctypes.POINTER(GppProblem))
    L.gpp_batch_create.restype = ctypes.c_int
    L.gpp_batch_create.argtypes = [I32, PP, ctypes.POINTER(Options), ctypes.POINTER(P)]
    L.gpp_batch_run.restype = ctypes.c_int
    L.gpp_batch_run.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(P), P, ctypes.POINTER(ctypes.c_float)]
    L.gpp_batch_config.restype = ctypes.c_int
    L.gpp_batch_config.argtypes = [P]
    L.gpp_batch_workspace_bytes.restype = SZ
    L.gpp_batch_workspace_bytes.argtypes = [P]
    L.gpp_batch_destroy.restype = None
    L.gpp_batch_destroy.argtypes = [P]
    L.gpp_kernel_batched.restype = ctypes.c_int
    L.gpp_kernel_batched.argtypes = [I32, PP, ctypes.POINTER(Options), ctypes.POINTER(P), ctypes.POINTER(P)]
    return L


lib = _load()


def last_error() -> str:
    return lib.fk_last_error().decode()


def _check(rc):
    if rc != FK_OK:
        raise FkError(rc, last_error())


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if _is_torch(x):
        if not x.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return x.data_ptr()
    if not (isinstance(x, np.ndarray) and x.flags["C_CONTIGUOUS"]):
        raise ValueError("array must be a C-contiguous numpy array")
    return x.ctypes.data


_PRECISIONS = {"fp64": FK_PREC_FP64, "mixed": FK_PREC_MIXED}
_NP = {"c128": np.complex128, "c64": np.complex64, "f64": np.float64}


def _precision(precision):
    if precision not in _PRECISIONS:
        raise ValueError(f"precision must be 'fp64' or 'mixed', not {precision!r}")
    return _PRECISIONS[precision]


def _bulk(where, arrays):
    """Check the bulk arrays of one call before the library sees them.

    arrays: (name, x, kind, numel) with kind "c128" / "c64" / "f64" and numel
    the number of elements the kernels read.  Every array must be a
    C-contiguous numpy array (host memory) or every one a contiguous torch
    tensor on the current ROCm device (device memory), of exactly that
    element type and with at least numel elements.  Returns True for device
    memory.  Raises TypeError / ValueError, so a wrong buffer never reaches
    a kernel."""
    kinds = {_is_torch(x) for _, x, _, _ in arrays if x is not None}
    if len(kinds) != 1:
        raise TypeError(f"{where}: bulk arrays must be all numpy (host) or all torch device tensors")
    on_dev = kinds.pop()
    dev = None
    if on_dev:
        import torch
        tdt = {"c128": torch.complex128, "c64": torch.complex64, "f64": torch.float64}
        dev = torch.cuda.current_device() if torch.cuda.is_available() else None
    for name, x, kind, numel in arrays:
        if x is None:
            raise ValueError(f"{where}: {name} is missing")
        if on_dev:
            if not x.is_cuda:
                raise TypeError(f"{where}: {name} is a CPU tensor; pass numpy arrays for host memory")
            if x.device.index != dev:
                raise ValueError(f"{where}: {name} is on {x.device}, the library runs on cuda:{dev}")
            if x.dtype != tdt[kind]:
                raise TypeError(f"{where}: {name} must be {tdt[kind]}, not {x.dtype}")
            if not x.is_contiguous():
                raise ValueError(f"{where}: {name} must be contiguous")
            n = x.numel()
        else:
            if not isinstance(x, np.ndarray):
                raise TypeError(f"{where}: {name} must be a numpy array or a torch device tensor")
            if x.dtype != _NP[kind]:
                raise TypeError(f"{where}: {name} must be {np.dtype(_NP[kind])}, not {x.dtype}")
            if not x.flags["C_CONTIGUOUS"]:
                raise ValueError(f"{where}: {name} must be C-contiguous")
            n = x.size
        if n < numel:
            raise ValueError(f"{where}: {name} has {n} elements, the problem reads {numel}")
    return on_dev


def _host_len(where, name, a, n):
    if a.shape[0] < n:
        raise ValueError(f"{where}: {name} has {a.shape[0]} entries, the problem needs {n}")


def _workspace(where, workspace, on_dev):
    """(pointer, bytes) of a caller workspace: a device tensor for device
    memory, on the current device."""
    if workspace is None:
        return None, 0
    if not _is_torch(workspace) or not workspace.is_cuda:
        raise TypeError(f"{where}: the workspace must be a torch device tensor")
    import torch
    if workspace.device.index != torch.cuda.current_device():
        raise ValueError(f"{where}: the workspace is on {workspace.device}")
    return _ptr(workspace), int(workspace.numel() * workspace.element_size())


def _stream_handle(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream  # torch.cuda.Stream


def default_options() -> Options:
    return lib.fk_default_options()


# --------------------------------------------------------------------------
# GPP
# --------------------------------------------------------------------------
def gpp_kernel(nb, nv, ncouls, ngpown, wx, aqsm, aqsn, I_eps, wtilde, vcoul, inv_igp_index, indinv, *,
               to1=1e-6, limittwo=0.25, limitone=2.5e5, occ=1.0, precision="fp64", mask=True,
               stream=None, workspace=None, config=0, time_kernel=False, device=-1, acht_n1=False,
               achstemp=False, semantics=FK_GPP_CANONICAL):
    """Run gpp_kernel.  Returns (achtemp, asxtemp, kernel_ms|None), and with
    acht_n1=True (achtemp, asxtemp, kernel_ms|None, acht_n1_loc[nb]).
    achstemp=True (through gpp_kernel_ex) appends the static-CH remainder
    achstemp (a one-element array / tensor) as the last item.
    semantics=FK_GPP_GPPSUM evaluates the gppsum = 1 loop of gppKerSeq.cpp
    (FP64 only; achtemp and asxtemp).

    numpy inputs -> host memory mode, outputs are numpy complex128 arrays.
    torch device tensors -> device mode, outputs are torch complex128 device
    tensors (valid once the stream has synchronised)."""
    wx = np.ascontiguousarray(wx, np.float64)
    inv = np.ascontiguousarray(inv_igp_index, np.int32)
    ind = np.ascontiguousarray(indinv, np.int32)
    nw = int(wx.shape[0])
    prec = _precision(precision)
    ck = "c64" if prec == FK_PREC_MIXED else "c128"
    nbC, GC = int(nb) * int(ncouls), int(ngpown) * int(ncouls)
    on_dev = _bulk("gpp_kernel", [("aqsm", aqsm, ck, nbC), ("aqsn", aqsn, ck, nbC), ("I_eps", I_eps, ck, GC),
                                  ("wtilde", wtilde, ck, GC), ("vcoul", vcoul, "f64", int(ncouls))])
    _host_len("gpp_kernel", "inv_igp_index", inv, int(ngpown))
    _host_len("gpp_kernel", "indinv", ind, int(ncouls) + 1)
    prob = GppProblem(int(nb), int(nv), int(ncouls), int(ngpown), nw, wx.ctypes.data, _ptr(aqsm), _ptr(aqsn),
                      _ptr(I_eps), _ptr(wtilde), _ptr(vcoul), inv.ctypes.data, ind.ctypes.data,
                      float(to1), float(limittwo), float(limitone), float(occ), int(semantics))
    opt = default_options()
    opt.memory = FK_MEM_DEVICE if on_dev else FK_MEM_HOST
    opt.precision = prec
    opt.mask = 1 if mask else 0
    opt.device = int(device)
    opt.stream = _stream_handle(stream)
    opt.workspace, opt.workspace_bytes = _workspace("gpp_kernel", workspace, on_dev)
    opt.config = int(config)
    ms = ctypes.c_float(-1.0)
    if time_kernel:
        opt.kernel_ms = ctypes.pointer(ms)
    if on_dev:
        import torch
        ach = torch.empty(nw, dtype=torch.complex128, device=aqsn.device)
        asx = torch.empty(nw, dtype=torch.complex128, device=aqsn.device)
        n1 = torch.empty(max(int(nb), 1), dtype=torch.complex128, device=aqsn.device) if acht_n1 else None
        achs = torch.empty(1, dtype=torch.complex128, device=aqsn.device) if achstemp else None
    else:
        ach = np.zeros(nw, np.complex128)
        asx = np.zeros(nw, np.complex128)
        n1 = np.zeros(max(int(nb), 1), np.complex128) if acht_n1 else None
        achs = np.zeros(1, np.complex128) if achstemp else None
    if achstemp:
        outs = GppOutputs(_ptr(ach), _ptr(asx), _ptr(n1), _ptr(achs))
        _check(lib.gpp_kernel_ex(ctypes.byref(prob), ctypes.byref(opt), ctypes.byref(outs)))
    else:
        _check(lib.gpp_kernel(ctypes.byref(prob), ctypes.byref(opt), _ptr(ach), _ptr(asx), _ptr(n1)))
    out = (ach, asx, (float(ms.value) if time_kernel else None))
    if acht_n1:
        out = out + (n1[:int(nb)],)
    return out + (achs,) if achstemp else out


def gpp_workspace_size(nb, nv, ncouls, ngpown, nfreq, precision="fp64", config=0, acht_n1=False):
    """Device scratch bytes of gpp_kernel on this shape; acht_n1=True: of a
    call that also returns acht_n1_loc (gpp_kernel_ex_workspace_size)."""
    wx = np.zeros(nfreq)
    dummy = 1  # non-NULL placeholders: only sizes are read
    inv = np.zeros(ngpown, np.int32)
    ind = np.zeros(ncouls + 1, np.int32)
    prob = GppProblem(nb, nv, ncouls, ngpown, nfreq, wx.ctypes.data, dummy, dummy, dummy, dummy, dummy,
                      inv.ctypes.data, ind.ctypes.data, 1e-6, 0.25, 2.5e5, 1.0)
    opt = default_options()
    opt.precision = _precision(precision)
    opt.config = config
    if acht_n1:
        outs = GppOutputs(dummy, dummy, dummy, None)
        return int(lib.gpp_kernel_ex_workspace_size(ctypes.byref(prob), ctypes.byref(opt), ctypes.byref(outs)))
    return int(lib.gpp_kernel_workspace_size(ctypes.byref(prob), ctypes.byref(opt)))


def gpp_problem(p, **kw):
    """gpp_kernel on an object with the oracle's GppProblem fields."""
    return gpp_kernel(p.nb, p.nv, p.ncouls, p.ngpown, p.wx, p.aqsm, p.aqsn, p.I_eps, p.wtilde, p.vcoul,
                      p.inv_igp_index, p.indinv, to1=p.to1, limittwo=p.limittwo, limitone=p.limitone,
                      occ=p.occ, **kw)


# --------------------------------------------------------------------------
# batched GPP
# --------------------------------------------------------------------------
class FfBatch:
    """A batch of FF contractions on the device (ff_batch_create).

    problems: sequence of dicts with the ff_kernel arguments: nb, ncouls,
    ngpown, nfreq, aqsm, aqsn, I_eps, vcoul (torch device tensors),
    inv_igp_index, indinv and optionally achs_freq.  config 0 lets the
    library time the precision's configurations on the batch."""

    def __init__(self, problems, precision="fp64", config=0):
        self._keep = []
        structs = []
        prec = _precision(precision)
        for k, p in enumerate(problems):
            inv = np.ascontiguousarray(p["inv_igp_index"], np.int32)
            ind = np.ascontiguousarray(p["indinv"], np.int32)
            where = f"FfBatch problem {k}"
            nbC = int(p["nb"]) * int(p["ncouls"])
            if not _bulk(where, [("aqsm", p["aqsm"], "c128", nbC), ("aqsn", p["aqsn"], "c128", nbC),
                                 ("I_eps", p["I_eps"], "c128", int(p["nfreq"]) * int(p["ngpown"]) * int(p["ncouls"])),
                                 ("vcoul", p["vcoul"], "f64", int(p["ncouls"]))]):
                raise TypeError(f"{where}: batches take torch device tensors")
            _host_len(where, "inv_igp_index", inv, int(p["ngpown"]))
            _host_len(where, "indinv", ind, int(p["ncouls"]) + 1)
            self._keep += [inv, ind, p]
            structs.append(FfProblem(int(p["nb"]), int(p["ncouls"]), int(p["ngpown"]), int(p["nfreq"]),
                                     _ptr(p["aqsm"]), _ptr(p["aqsn"]), _ptr(p["I_eps"]), _ptr(p["vcoul"]),
                                     inv.ctypes.data, ind.ctypes.data, int(p.get("achs_freq", 0))))
        self.n = len(structs)
        self.shapes = [(s.number_bands, s.nfreq) for s in structs]
        self._structs = structs
        arr = (ctypes.POINTER(FfProblem) * self.n)(*[ctypes.pointer(s) for s in structs])
        opt = default_options()
        opt.memory = FK_MEM_DEVICE
        opt.precision = prec
        opt.config = int(config)
        h = ctypes.c_void_p()
        _check(lib.ff_batch_create(self.n, arr, ctypes.byref(opt), ctypes.byref(h)))
        self._h = h

    @property
    def config(self):
        return lib.ff_batch_config(self._h)

    @property
    def workspace_bytes(self):
        return int(lib.ff_batch_workspace_bytes(self._h))

    def run(self, stream=None, time_it=False, want_schDv=True, want_achs=True):
        """Returns ([schDt_i], [schDv_i] | None, [achsDtemp_i] | None, ms | None)
        as torch device tensors."""
        import torch
        mk = lambda *s: torch.empty(*s, dtype=torch.complex128, device="cuda")  # noqa: E731
        t = [mk(nb, nf) for nb, nf in self.shapes]
        v = [mk(nb, nf) for nb, nf in self.shapes] if want_schDv else None
        a = [mk(1) for _ in self.shapes] if want_achs else None
        arr = lambda xs: (ctypes.c_void_p * self.n)(*[x.data_ptr() for x in xs]) if xs is not None else None  # noqa: E731
        ms = ctypes.c_float(-1.0)
        _check(lib.ff_batch_run(self._h, arr(t), arr(v), arr(a), _stream_handle(stream),
                                ctypes.byref(ms) if time_it else None))
        return t, v, a, (float(ms.value) if time_it else None)

    def close(self):
        if getattr(self, "_h", None):
            lib.ff_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class GppBatch:
    """A batch of GPP problems on the device (gpp_batch_create).

    problems: sequence of dicts with the gpp_kernel arguments: nb, nv, ncouls,
    ngpown, wx, aqsm, aqsn, I_eps, wtilde, vcoul (torch device tensors),
    inv_igp_index, indinv and optionally to1, limittwo, limitone, occ.
    FP64: config 0 lets the library time every tile configuration on the
    batch.  Mixed (complex64 tensors): config 0 runs the int8-split
    contraction of gpp_kernel's mixed path, config 1..4 the tile kernel with
    FP32 MFMA."""

    def __init__(self, problems, precision="fp64", config=0, mask=True):
        self._keep = []
        structs = []
        prec = _precision(precision)
        ck = "c64" if prec == FK_PREC_MIXED else "c128"
        for k, p in enumerate(problems):
            wx = np.ascontiguousarray(p["wx"], np.float64)
            inv = np.ascontiguousarray(p["inv_igp_index"], np.int32)
            ind = np.ascontiguousarray(p["indinv"], np.int32)
            where = f"GppBatch problem {k}"
            nbC, GC = int(p["nb"]) * int(p["ncouls"]), int(p["ngpown"]) * int(p["ncouls"])
            if not _bulk(where, [("aqsm", p["aqsm"], ck, nbC), ("aqsn", p["aqsn"], ck, nbC),
                                 ("I_eps", p["I_eps"], ck, GC), ("wtilde", p["wtilde"], ck, GC),
                                 ("vcoul", p["vcoul"], "f64", int(p["ncouls"]))]):
                raise TypeError(f"{where}: batches take torch device tensors")
            _host_len(where, "inv_igp_index", inv, int(p["ngpown"]))
            _host_len(where, "indinv", ind, int(p["ncouls"]) + 1)
            self._keep += [wx, inv, ind, p]
            structs.append(GppProblem(int(p["nb"]), int(p["nv"]), int(p["ncouls"]), int(p["ngpown"]), int(wx.shape[0]),
                                      wx.ctypes.data, _ptr(p["aqsm"]), _ptr(p["aqsn"]), _ptr(p["I_eps"]),
                                      _ptr(p["wtilde"]), _ptr(p["vcoul"]), inv.ctypes.data, ind.ctypes.data,
                                      float(p.get("to1", 1e-6)), float(p.get("limittwo", 0.25)),
                                      float(p.get("limitone", 2.5e5)), float(p.get("occ", 1.0))))
        self.n = len(structs)
        self.nw = [s.nfreq for s in structs]
        self._structs = structs
        arr = (ctypes.POINTER(GppProblem) * self.n)(*[ctypes.pointer(s) for s in structs])
        opt = default_options()
        opt.memory = FK_MEM_DEVICE
        opt.precision = prec
        opt.mask = 1 if mask else 0
        opt.config = int(config)
        h = ctypes.c_void_p()
        _check(lib.gpp_batch_create(self.n, arr, ctypes.byref(opt), ctypes.byref(h)))
        self._h = h

    @property
    def config(self):
        return lib.gpp_batch_config(self._h)

    @property
    def workspace_bytes(self):
        return int(lib.gpp_batch_workspace_bytes(self._h))

    def run(self, stream=None, time_it=False, out=None):
        """Returns ([achtemp_i], [asxtemp_i], ms|None) as torch device tensors."""
        import torch
        if out is None:
            out = ([torch.empty(w, dtype=torch.complex128, device="cuda") for w in self.nw],
                   [torch.empty(w, dtype=torch.complex128, device="cuda") for w in self.nw])
        ach, asx = out
        pa = (ctypes.c_void_p * self.n)(*[t.data_ptr() for t in ach])
        px = (ctypes.c_void_p * self.n)(*[t.data_ptr() for t in asx])
        ms = ctypes.c_float(-1.0)
        _check(lib.gpp_batch_run(self._h, pa, px, _stream_handle(stream), ctypes.byref(ms) if time_it else None))
        return ach, asx, (float(ms.value) if time_it else None)

    def close(self):
        if getattr(self, "_h", None):
            lib.gpp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


# --------------------------------------------------------------------------
# FF
# --------------------------------------------------------------------------
def ff_kernel(nb, ncouls, ngpown, nfreq, aqsm, aqsn, I_eps, vcoul, inv_igp_index, indinv, *, achs_freq=0,
              stream=None, workspace=None, config=0, time_kernel=False, want_schDv=True, device=-1,
              precision="fp64"):
    """Run ff_kernel.  Returns (schDt[nb][nfreq], schDv or None, achsDtemp, kernel_ms|None).
    precision "mixed" takes the same complex128 inputs and runs the contraction
    in split BF16 on MFMA: operands rounded to FP32, FP32 sums per 1024-ig
    chunk, the chunk sums combined in FP64 (FP32-class accuracy, ~7e-7)."""
    inv = np.ascontiguousarray(inv_igp_index, np.int32)
    ind = np.ascontiguousarray(indinv, np.int32)
    prec = _precision(precision)
    nbC = int(nb) * int(ncouls)
    on_dev = _bulk("ff_kernel", [("aqsm", aqsm, "c128", nbC), ("aqsn", aqsn, "c128", nbC),
                                 ("I_eps", I_eps, "c128", int(nfreq) * int(ngpown) * int(ncouls)),
                                 ("vcoul", vcoul, "f64", int(ncouls))])
    _host_len("ff_kernel", "inv_igp_index", inv, int(ngpown))
    _host_len("ff_kernel", "indinv", ind, int(ncouls) + 1)
    prob = FfProblem(int(nb), int(ncouls), int(ngpown), int(nfreq), _ptr(aqsm), _ptr(aqsn), _ptr(I_eps),
                     _ptr(vcoul), inv.ctypes.data, ind.ctypes.data, int(achs_freq))
    opt = default_options()
    opt.memory = FK_MEM_DEVICE if on_dev else FK_MEM_HOST
    opt.precision = prec
    opt.device = int(device)
    opt.stream = _stream_handle(stream)
    opt.workspace, opt.workspace_bytes = _workspace("ff_kernel", workspace, on_dev)
    opt.config = int(config)
    ms = ctypes.c_float(-1.0)
    if time_kernel:
        opt.kernel_ms = ctypes.pointer(ms)
    if on_dev:
        import torch
        mk = lambda *s: torch.empty(*s, dtype=torch.complex128, device=aqsn.device)  # noqa: E731
    else:
        mk = lambda *s: np.zeros(s, np.complex128)  # noqa: E731
    schDt = mk(nb, nfreq)
    schDv = mk(nb, nfreq) if want_schDv else None
    achs = mk(1)
    _check(lib.ff_kernel(ctypes.byref(prob), ctypes.byref(opt), _ptr(schDt), _ptr(schDv), _ptr(achs)))
    return schDt, schDv, achs, (float(ms.value) if time_kernel else None)


def ff_workspace_size(nb, ncouls, ngpown, nfreq, config=0, precision="fp64"):
    dummy = 1  # non-NULL placeholders: only sizes are read
    inv = np.zeros(ngpown, np.int32)
    ind = np.zeros(ncouls + 1, np.int32)
    prob = FfProblem(nb, ncouls, ngpown, nfreq, dummy, dummy, dummy, dummy, inv.ctypes.data, ind.ctypes.data, 0)
    opt = default_options()
    opt.precision = _precision(precision)
    opt.config = config
    return int(lib.ff_kernel_workspace_size(ctypes.byref(prob), ctypes.byref(opt)))


def ff_problem(p, **kw):
    return ff_kernel(p.nb, p.ncouls, p.ngpown, p.nfreq, p.aqsm, p.aqsn, p.I_eps, p.vcoul, p.inv_igp_index,
                     p.indinv, **kw)


def _ff_full_struct(nb, nv, ncouls, ngpown, nfreq, nfreqeval, aqsm, aqsn, I_epsR, I_epsA, vcoul, host, scal):
    return FfFullProblem(int(nb), int(nv), int(ncouls), int(ngpown), int(nfreq), int(nfreqeval), _ptr(aqsm),
                         _ptr(aqsn), _ptr(I_epsR), _ptr(I_epsA), _ptr(vcoul), *[h.ctypes.data for h in host],
                         *[float(x) for x in scal[:4]], int(scal[4]), int(scal[5]) if len(scal) > 5 else 0)


def ff_kernel_full(nb, nv, ncouls, ngpown, nfreq, nfreqeval, aqsm, aqsn, I_epsR, I_epsA, vcoul, inv_igp_index,
                   indinv, ekq, dFreqGrid, pref, dFreqBrd, *, pref_zb=0.5 / 3.14, freqevalmin=0.0,
                   freqevalstep=0.5, occ=1.0, achs_freq=0, semantics=FK_FF_FORTRAN, stream=None, workspace=None,
                   config=0, time_kernel=False, want_schDt=True, device=-1, precision="fp64"):
    """Run ff_kernel_full (the whole FF evaluation of the Fortran program, or
    with semantics FK_FF_CPP_PORT[_ZERO_SCHDT] that of the C++ port).
    Bulk arrays are numpy (host) or torch device tensors; inv_igp_index,
    indinv, ekq, dFreqGrid, pref, dFreqBrd are host arrays.  Returns a dict
    with achsDtemp, asxDtemp, achDtemp, achDtemp_cor, achDtemp_corb,
    ach2Dtemp, schDt (or None) as numpy / torch like the inputs, and
    kernel_ms (the call's device time, or None)."""
    host = [np.ascontiguousarray(inv_igp_index, np.int32), np.ascontiguousarray(indinv, np.int32),
            np.ascontiguousarray(ekq, np.float64), np.ascontiguousarray(dFreqGrid, np.float64),
            np.ascontiguousarray(pref, np.float64), np.ascontiguousarray(dFreqBrd, np.complex128)]
    prec = _precision(precision)
    nbC, fGC = int(nb) * int(ncouls), int(nfreq) * int(ngpown) * int(ncouls)
    on_dev = _bulk("ff_kernel_full", [("aqsm", aqsm, "c128", nbC), ("aqsn", aqsn, "c128", nbC),
                                      ("I_epsR", I_epsR, "c128", fGC), ("I_epsA", I_epsA, "c128", fGC),
                                      ("vcoul", vcoul, "f64", int(ncouls))])
    for name, a, n in zip(("inv_igp_index", "indinv", "ekq", "dFreqGrid", "pref", "dFreqBrd"), host,
                          (ngpown, ncouls + 1, nb, nfreq, nfreq, nfreq)):
        _host_len("ff_kernel_full", name, a, int(n))
    prob = _ff_full_struct(nb, nv, ncouls, ngpown, nfreq, nfreqeval, aqsm, aqsn, I_epsR, I_epsA, vcoul, host,
                           (pref_zb, freqevalmin, freqevalstep, occ, achs_freq, semantics))
    opt = default_options()
    opt.memory = FK_MEM_DEVICE if on_dev else FK_MEM_HOST
    opt.precision = prec
    opt.device = int(device)
    opt.stream = _stream_handle(stream)
    opt.workspace, opt.workspace_bytes = _workspace("ff_kernel_full", workspace, on_dev)
    opt.config = int(config)
    ms = ctypes.c_float(-1.0)
    if time_kernel:
        opt.kernel_ms = ctypes.pointer(ms)
    if on_dev:
        import torch
        mk = lambda *s: torch.zeros(*s, dtype=torch.complex128, device=aqsn.device)  # noqa: E731
    else:
        mk = lambda *s: np.zeros(s, np.complex128)  # noqa: E731
    res = {"achsDtemp": mk(1)}
    for k in FF_FULL_OUTPUTS:
        res[k] = mk(nfreqeval)
    res["schDt"] = mk(nb, nfreq) if want_schDt else None
    outs = FfFullOutputs(*[_ptr(res[k]) for k in ("achsDtemp",) + FF_FULL_OUTPUTS + ("schDt",)])
    _check(lib.ff_kernel_full(ctypes.byref(prob), ctypes.byref(opt), ctypes.byref(outs)))
    res["kernel_ms"] = float(ms.value) if time_kernel else None
    return res


def ff_full_workspace_size(nb, nv, ncouls, ngpown, nfreq, nfreqeval, ekq, dFreqGrid, config=0, precision="fp64",
                           freqevalmin=0.0, freqevalstep=0.5, semantics=FK_FF_FORTRAN):
    """Workspace bytes of ff_kernel_full; it depends on the evaluation axis
    and the semantics (the I_epsA bands and frequencies that are read)."""
    dummy = np.zeros(1)
    host = [np.zeros(ngpown, np.int32), np.zeros(ncouls + 1, np.int32), np.ascontiguousarray(ekq, np.float64),
            np.ascontiguousarray(dFreqGrid, np.float64), np.zeros(nfreq), np.zeros(nfreq, np.complex128)]
    prob = _ff_full_struct(nb, nv, ncouls, ngpown, nfreq, nfreqeval, dummy, dummy, dummy, dummy, dummy, host,
                           (0.0, freqevalmin, freqevalstep, 1.0, 0, semantics))
    opt = default_options()
    opt.precision = _precision(precision)
    opt.config = config
    return int(lib.ff_kernel_full_workspace_size(ctypes.byref(prob), ctypes.byref(opt)))


def ff_full_problem(p, **kw):
    """ff_kernel_full on an object with the oracle's FfFullProblem fields."""
    return ff_kernel_full(p.nb, p.nv, p.ncouls, p.ngpown, p.nfreq, p.nfreqeval, p.aqsm, p.aqsn, p.I_epsR,
                          p.I_epsA, p.vcoul, p.inv_igp_index, p.indinv, p.ekq, p.dFreqGrid, p.pref, p.dFreqBrd,
                          pref_zb=p.pref_zb, freqevalmin=p.freqevalmin, freqevalstep=p.freqevalstep, occ=p.occ,
                          achs_freq=p.achs_freq, **kw)


# --------------------------------------------------------------------------
# instance files (fk_io.cpp)
# --------------------------------------------------------------------------
class _Fields:
    """Plain attribute bag for the arrays and sizes of an instance."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def gpp_instance_write(path, p, complex64=False):
    """Write a GPP problem (an object with the oracle GppProblem's fields,
    numpy arrays) to an instance file."""
    cdt = np.complex64 if complex64 else np.complex128
    keep = [np.ascontiguousarray(p.wx, np.float64), np.ascontiguousarray(p.inv_igp_index, np.int32),
            np.ascontiguousarray(p.indinv, np.int32), np.ascontiguousarray(p.vcoul, np.float64)]
    cx = [np.ascontiguousarray(getattr(p, k), cdt) for k in ("aqsm", "aqsn", "I_eps", "wtilde")]
    prob = GppProblem(int(p.nb), int(p.nv), int(p.ncouls), int(p.ngpown), len(keep[0]), keep[0].ctypes.data,
                      cx[0].ctypes.data, cx[1].ctypes.data, cx[2].ctypes.data, cx[3].ctypes.data,
                      keep[3].ctypes.data, keep[1].ctypes.data, keep[2].ctypes.data, float(p.to1),
                      float(p.limittwo), float(p.limitone), float(p.occ))
    _check(lib.fk_gpp_instance_write(os.fsencode(path), ctypes.byref(prob), 1 if complex64 else 0))


def ff_instance_write(path, p):
    """Write an FF evaluation problem (the oracle FfFullProblem's fields)."""
    host = [np.ascontiguousarray(p.inv_igp_index, np.int32), np.ascontiguousarray(p.indinv, np.int32),
            np.ascontiguousarray(p.ekq, np.float64), np.ascontiguousarray(p.dFreqGrid, np.float64),
            np.ascontiguousarray(p.pref, np.float64), np.ascontiguousarray(p.dFreqBrd, np.complex128)]
    cx = [np.ascontiguousarray(getattr(p, k), np.complex128) for k in ("aqsm", "aqsn", "I_epsR", "I_epsA")]
    v = np.ascontiguousarray(p.vcoul, np.float64)
    prob = _ff_full_struct(p.nb, p.nv, p.ncouls, p.ngpown, p.nfreq, p.nfreqeval, *cx, v, host,
                           (p.pref_zb, p.freqevalmin, p.freqevalstep, p.occ, p.achs_freq))
    _check(lib.fk_ff_instance_write(os.fsencode(path), ctypes.byref(prob)))


def _view(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    buf = (ctypes.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype, count=n)


class Instance:
    """An instance file opened with fk_instance_open: .kind ("gpp" / "ff"),
    .problem (the arrays as read-only numpy views of the mapping, with the
    oracle problem classes' field names) and .raw (the C problem struct).
    Close it (or use it as a context manager) only after the last use of
    the views."""

    def __init__(self, path):
        h = ctypes.c_void_p()
        _check(lib.fk_instance_open(os.fsencode(path), ctypes.byref(h)))
        self._h = h
        kind = lib.fk_instance_kind(h)
        if kind == 1:
            self.kind = "gpp"
            q = lib.fk_instance_gpp(h).contents
            cdt = np.complex64 if lib.fk_instance_complex_bytes(h) == 8 else np.complex128
            nb, C, G = q.number_bands, q.ncouls, q.ngpown
            self.problem = _Fields(
                nb=nb, nv=q.nvband, ncouls=C, ngpown=G, wx=_view(q.wx, q.nfreq, np.float64),
                aqsm=_view(q.aqsmtemp, nb * C, cdt).reshape(nb, C), aqsn=_view(q.aqsntemp, nb * C, cdt).reshape(nb, C),
                I_eps=_view(q.I_eps_array, G * C, cdt).reshape(G, C),
                wtilde=_view(q.wtilde_array, G * C, cdt).reshape(G, C), vcoul=_view(q.vcoul, C, np.float64),
                inv_igp_index=_view(q.inv_igp_index, G, np.int32), indinv=_view(q.indinv, C + 1, np.int32),
                to1=q.to1, limittwo=q.limittwo, limitone=q.limitone, occ=q.occ)
            self.raw = q
        else:
            self.kind = "ff"
            q = lib.fk_instance_ff(h).contents
            nb, C, G, nf = q.number_bands, q.ncouls, q.ngpown, q.nfreq
            self.problem = _Fields(
                nb=nb, nv=q.nvband, ncouls=C, ngpown=G, nfreq=nf, nfreqeval=q.nfreqeval,
                aqsm=_view(q.aqsmtemp, nb * C, np.complex128).reshape(nb, C),
                aqsn=_view(q.aqsntemp, nb * C, np.complex128).reshape(nb, C),
                I_epsR=_view(q.I_epsR_array, nf * G * C, np.complex128).reshape(nf, G, C),
                I_epsA=_view(q.I_epsA_array, nf * G * C, np.complex128).reshape(nf, G, C),
                vcoul=_view(q.vcoul, C, np.float64), inv_igp_index=_view(q.inv_igp_index, G, np.int32),
                indinv=_view(q.indinv, C + 1, np.int32), ekq=_view(q.ekq, nb, np.float64),
                dFreqGrid=_view(q.dFreqGrid, nf, np.float64), pref=_view(q.pref, nf, np.float64),
                dFreqBrd=_view(q.dFreqBrd, nf, np.complex128), pref_zb=q.pref_zb, freqevalmin=q.freqevalmin,
                freqevalstep=q.freqevalstep, occ=q.occ, achs_freq=q.achs_freq)
            self.raw = q

    def close(self):
        if self._h is not None:
            lib.fk_instance_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def instance_list_read(path):
    arr = ctypes.POINTER(ctypes.c_char_p)()
    n = ctypes.c_int32(0)
    _check(lib.fk_instance_list_read(os.fsencode(path), ctypes.byref(arr), ctypes.byref(n)))
    try:
        return [os.fsdecode(arr[i]) for i in range(n.value)]
    finally:
        lib.fk_instance_list_free(arr, n)


# --------------------------------------------------------------------------
# device fill / complex helpers / diagnostics
# --------------------------------------------------------------------------
def fill_complex(dst, seed, stream_id, re, im, stream=None):
    """Counter-based fill of a complex128 or complex64 device tensor."""
    import torch
    if not (_is_torch(dst) and dst.is_cuda and dst.is_contiguous()
            and dst.dtype in (torch.complex128, torch.complex64)):
        raise TypeError("fill_complex: dst must be a contiguous complex128 / complex64 device tensor")
    as64 = 1 if dst.dtype == torch.complex64 else 0
    _check(lib.fk_fill_complex(_ptr(dst), dst.numel(), seed, stream_id, re[0], re[1], im[0], im[1], as64,
                               _stream_handle(stream)))


def fill_real(dst, seed, stream_id, lo, hi, stream=None):
    import torch
    if not (_is_torch(dst) and dst.is_cuda and dst.is_contiguous() and dst.dtype == torch.float64):
        raise TypeError("fill_real: dst must be a contiguous float64 device tensor")
    _check(lib.fk_fill_real(_ptr(dst), dst.numel(), seed, stream_id, lo, hi, _stream_handle(stream)))


def complex_apply(op, a=None, b=None, s=None, out=None, on_device=False):
    n = len(next(x for x in (a, b, s, out) if x is not None))
    a = None if a is None else np.ascontiguousarray(a, np.complex128)
    b = None if b is None else np.ascontiguousarray(b, np.complex128)
    s = None if s is None else np.ascontiguousarray(s, np.float64)
    out = np.zeros(n, np.complex128) if out is None else np.array(out, np.complex128)
    _check(lib.fk_complex_apply(COP[op], n, _ptr(a), _ptr(b), _ptr(s), _ptr(out), int(on_device)))
    return out


def mfma_probe(kind):
    """kind 0: FP64 16x16x4 MFMA, kind 1: FP32 16x16x4 MFMA; returns [64 lanes][4 registers]."""
    out = np.zeros(256, np.float64)
    _check(lib.fk_mfma_probe(int(kind), out.ctypes.data))
    return out.reshape(64, 4)
