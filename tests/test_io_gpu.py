"""Instance files on the GPU: problems read back from files (the library's
mapping, the host programs' --input / --dump and the batch of an instance
list) give the oracle's results.  FP64 bar 1e-12; the int8-split mixed form
1e-8; the batched mixed form (FP32 MFMA) 1e-5."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "fortkernels-_amd", "bin")


def _run(exe, *args, timeout=180):
    print(f"{exe} {' '.join(map(str, args))}", flush=True)
    r = subprocess.run([os.path.join(BIN, exe), *map(str, args)], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def _lines(text, name, prefix=""):
    out = {}
    for m in re.finditer(rf"^{prefix}{name}\[(\d+)\] = \(([^,]+),([^)]+)\)$", text, re.M):
        out[int(m.group(1))] = complex(float(m.group(2)), float(m.group(3)))
    return np.array([out[i] for i in sorted(out)])


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def test_gpp_instance_library_and_cli(fk, fko, tmp_path):
    p = fko.gpp_random(40, 4, 600, 30, nw=5, seed=12)
    path = str(tmp_path / "g.fki")
    fk.gpp_instance_write(path, p)
    ref_a, ref_s = fko.gpp(p)[:2]
    with fk.Instance(path) as inst:
        q = inst.problem
        ach, asx, _ = fk.gpp_kernel(q.nb, q.nv, q.ncouls, q.ngpown, q.wx, q.aqsm, q.aqsn, q.I_eps, q.wtilde, q.vcoul,
                                    q.inv_igp_index, q.indinv, to1=q.to1, limittwo=q.limittwo,
                                    limitone=q.limitone, occ=q.occ)
    assert _rel(ach, ref_a) <= 1e-12 and _rel(asx, ref_s) <= 1e-12
    out = _run("gppKer", "--input", path, "--achstemp")
    assert _rel(_lines(out, "achtemp"), ref_a) <= 1e-12 and _rel(_lines(out, "asxtemp"), ref_s) <= 1e-12
    m = re.search(r"^achstemp = \(([^,]+),([^)]+)\)$", out, re.M)
    want = fko.gpp_achstemp(p)
    assert abs(complex(float(m.group(1)), float(m.group(2))) - want) <= 1e-12 * abs(want)


def test_gpp_complex64_instance_runs_mixed(fk, fko, tmp_path):
    p = fko.gpp_random(64, 2, 1024, 51, seed=13)
    path = str(tmp_path / "g64.fki")
    fk.gpp_instance_write(path, p, complex64=True)
    for k in ("aqsm", "aqsn", "I_eps", "wtilde"):
        setattr(p, k, getattr(p, k).astype(np.complex64).astype(np.complex128))
    ref_a, ref_s = fko.gpp(p)[:2]
    out = _run("gppKer", "--input", path)
    assert "mixed" in out
    assert _rel(_lines(out, "achtemp"), ref_a) <= 1e-8 and _rel(_lines(out, "asxtemp"), ref_s) <= 1e-8


def test_gppker_dump_then_input_same_bits(fk, fko, tmp_path):
    path = str(tmp_path / "dump.fki")
    a = _run("gppKer", 16, 2, 512, 20, "--random", 5, "--dump", path)
    b = _run("gppKer", "--input", path)
    for name in ("achtemp", "asxtemp"):
        assert np.array_equal(_lines(a, name), _lines(b, name)), name
    with fk.Instance(path) as inst:        # the dumped arrays are the seeded generator's
        q = inst.problem
        p = fko.gpp_random(16, 2, 512, 25, seed=5)
        assert np.array_equal(q.aqsm, p.aqsm) and np.array_equal(q.wtilde, p.wtilde)
        assert np.array_equal(q.inv_igp_index, p.inv_igp_index)


def test_ff_instance_cli(fk, fko, tmp_path):
    p = fko.ff_full_random(30, 6, 90, 20, 7, 9, seed=14)
    p.freqevalmin = 0.35
    path = str(tmp_path / "f.fki")
    fk.ff_instance_write(path, p)
    ref = fko.ff_full(p)
    out = _run("ffKer", "--input", path)
    for name in fko.FF_OUTPUTS:
        assert _rel(_lines(out, name), ref[name]) <= 1e-12, name
    dump = str(tmp_path / "fd.fki")
    a = _run("ffKer", 12, 3, 40, 10, 5, 6, "--random", 9, "--achs-freq", 0, "--dump", dump)
    b = _run("ffKer", "--input", dump)
    for name in fko.FF_OUTPUTS:
        assert np.array_equal(_lines(a, name), _lines(b, name)), name
    # the C++ port's semantics on the same instance (the file holds inputs;
    # the evaluation is chosen on the command line)
    out = _run("ffKer", "--input", path, "--semantics", 1)
    q = fko.ff_full_random(30, 6, 90, 20, 7, 9, seed=14)
    q.freqevalmin = 0.35
    ref1 = fko.ff_full(q, variant=1, schDt_in=ref["schDt"])
    for name in fko.FF_OUTPUTS:
        got = _lines(out, name)
        if np.max(np.abs(ref1[name])) == 0:
            assert np.max(np.abs(got)) == 0, name
        else:
            assert _rel(got, ref1[name]) <= 1e-12, name


@pytest.mark.parametrize("complex64", [False, True])
def test_gppbatch_instance_list(fk, fko, tmp_path, complex64):
    shapes = [(48, 3, 640, 32, 3), (20, 2, 256, 12, 5), (70, 8, 1000, 50, 2)]
    names = []
    refs = []
    for i, (nb, nv, C, G, nw) in enumerate(shapes):
        p = fko.gpp_random(nb, nv, C, G, nw=nw, seed=30 + i)
        name = f"p{i}.fki"
        fk.gpp_instance_write(str(tmp_path / name), p, complex64=complex64)
        if complex64:
            for k in ("aqsm", "aqsn", "I_eps", "wtilde"):
                setattr(p, k, getattr(p, k).astype(np.complex64).astype(np.complex128))
        refs.append(fko.gpp(p)[:2])
        names.append(name)
    lst = tmp_path / "batch.txt"
    lst.write_text("# three problems\n" + "\n".join(names) + "\n")
    out = _run("gppBatch", str(lst), "--config", 2)
    tol = 1e-5 if complex64 else 1e-12
    for i, (ra, rs) in enumerate(refs):
        ga, gs = _lines(out, "achtemp", f"problem {i} "), _lines(out, "asxtemp", f"problem {i} ")
        assert _rel(ga, ra) <= tol and _rel(gs, rs) <= tol, (i, _rel(ga, ra), _rel(gs, rs))
    assert "Batch time" in out
