"""Instance files (fk_io.cpp through the C ABI; no GPU needed): round trips
of GPP (complex128 and complex64) and FF problems, the documented byte
layout read back independently with struct/numpy, rejection of damaged
files, and instance lists."""
import os
import struct

import numpy as np
import pytest

HEADER = struct.Struct("<8sii8i8d16q")


def _parse(path):
    raw = open(path, "rb").read()
    magic, kind, cbytes, *rest = HEADER.unpack_from(raw, 0)
    sizes, scalars, offsets = tuple(rest[:8]), tuple(rest[8:16]), tuple(rest[16:])
    return raw, magic, kind, cbytes, sizes, scalars, offsets


def _arr(raw, off, dtype, n):
    return np.frombuffer(raw, dtype=dtype, count=n, offset=off)


@pytest.mark.parametrize("complex64", [False, True])
def test_gpp_round_trip_and_layout(fk, fko, tmp_path, complex64):
    p = fko.gpp_random(9, 3, 70, 11, nw=4, seed=2)
    path = str(tmp_path / "g.fki")
    fk.gpp_instance_write(path, p, complex64=complex64)
    cdt = np.complex64 if complex64 else np.complex128
    raw, magic, kind, cbytes, sizes, scalars, offsets = _parse(path)
    assert magic == b"FKINST01" and kind == 1 and cbytes == (8 if complex64 else 16)
    assert sizes[:5] == (9, 3, 70, 11, 4) and sizes[5:] == (0, 0, 0)
    assert scalars[:4] == (p.to1, p.limittwo, p.limitone, p.occ)
    assert all(o % 64 == 0 for o in offsets[:8]) and offsets[0] == 256 and offsets[8:] == (0,) * 8
    assert np.array_equal(_arr(raw, offsets[0], "<f8", 4), p.wx)
    assert np.array_equal(_arr(raw, offsets[1], "<i4", 11), p.inv_igp_index)
    assert np.array_equal(_arr(raw, offsets[2], "<i4", 71), p.indinv)
    assert np.array_equal(_arr(raw, offsets[3], "<f8", 70), p.vcoul)
    for i, k, n in ((4, "aqsm", 9 * 70), (5, "aqsn", 9 * 70), (6, "I_eps", 11 * 70), (7, "wtilde", 11 * 70)):
        assert np.array_equal(_arr(raw, offsets[i], cdt, n), getattr(p, k).astype(cdt).reshape(-1)), k
    with fk.Instance(path) as inst:
        q = inst.problem
        assert inst.kind == "gpp" and (q.nb, q.nv, q.ncouls, q.ngpown) == (9, 3, 70, 11)
        for k in ("aqsm", "aqsn", "I_eps", "wtilde"):
            assert q.__dict__[k].dtype == cdt and np.array_equal(q.__dict__[k], getattr(p, k).astype(cdt)), k
        for k in ("wx", "vcoul", "inv_igp_index", "indinv"):
            assert np.array_equal(q.__dict__[k], getattr(p, k)), k
        assert (q.to1, q.limittwo, q.limitone, q.occ) == (p.to1, p.limittwo, p.limitone, p.occ)
        if not complex64:   # the oracle on the mapped arrays is the oracle on the originals
            a = fko.gpp(p)
            b = fko.gpp(fko.GppProblem(q.nb, q.nv, q.ncouls, q.ngpown, q.wx, q.aqsm, q.aqsn, q.I_eps, q.wtilde,
                                       q.vcoul, q.inv_igp_index, q.indinv))
            assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_ff_round_trip_and_layout(fk, fko, tmp_path):
    p = fko.ff_full_random(6, 2, 24, 5, 4, 3, seed=8)
    p.achs_freq = 2
    p.freqevalmin = -0.75
    path = str(tmp_path / "f.fki")
    fk.ff_instance_write(path, p)
    raw, magic, kind, cbytes, sizes, scalars, offsets = _parse(path)
    assert magic == b"FKINST01" and kind == 2 and cbytes == 16
    assert sizes[:7] == (6, 2, 24, 5, 4, 3, 2)
    assert scalars[:4] == (p.pref_zb, -0.75, p.freqevalstep, p.occ)
    assert np.array_equal(_arr(raw, offsets[0], "<f8", 6), p.ekq)
    assert np.array_equal(_arr(raw, offsets[3], "<c16", 4), p.dFreqBrd)
    assert np.array_equal(_arr(raw, offsets[10], "<c16", 4 * 5 * 24), p.I_epsA.reshape(-1))
    assert len(raw) >= offsets[10] + 16 * 4 * 5 * 24
    with fk.Instance(path) as inst:
        q = inst.problem
        assert inst.kind == "ff" and (q.nfreq, q.nfreqeval, q.achs_freq) == (4, 3, 2)
        for k in ("aqsm", "aqsn", "I_epsR", "I_epsA", "vcoul", "ekq", "dFreqGrid", "pref", "dFreqBrd",
                  "inv_igp_index", "indinv"):
            assert np.array_equal(q.__dict__[k], getattr(p, k)), k
        assert (q.pref_zb, q.freqevalmin, q.freqevalstep, q.occ) == (p.pref_zb, -0.75, p.freqevalstep, p.occ)


def test_damaged_files_rejected(fk, fko, tmp_path):
    p = fko.gpp_random(4, 1, 32, 4, seed=1)
    good = str(tmp_path / "good.fki")
    fk.gpp_instance_write(good, p)
    raw = bytearray(open(good, "rb").read())
    cases = {}
    bad = bytearray(raw); bad[0:8] = b"NOTANINS"; cases["magic"] = bad
    cases["truncated"] = raw[: len(raw) - 100]
    bad = bytearray(raw); struct.pack_into("<q", bad, 8 + 8 + 32 + 64 + 8 * 7, len(raw) + 64); cases["offset"] = bad
    bad = bytearray(raw); struct.pack_into("<i", bad, 16, -5); cases["size"] = bad
    bad = bytearray(raw); struct.pack_into("<i", bad, 8, 7); cases["kind"] = bad
    cases["short"] = raw[:100]
    for name, data in cases.items():
        path = str(tmp_path / f"{name}.fki")
        open(path, "wb").write(bytes(data))
        with pytest.raises(fk.FkError) as e:
            fk.Instance(path)
        assert e.value.code == fk.FK_ERR_ARG, name
        assert "fk_instance_open" in fk.last_error(), name
    with pytest.raises(fk.FkError):
        fk.Instance(str(tmp_path / "missing.fki"))


def test_instance_list(fk, tmp_path):
    sub = tmp_path / "lists"
    sub.mkdir()
    lst = sub / "batch.txt"
    lst.write_text("# three problems\na.fki\n\n  ../b.fki  \n/abs/c.fki\r\n# done\n")
    got = fk.instance_list_read(str(lst))
    assert got == [str(sub / "a.fki"), str(sub / "../b.fki"), "/abs/c.fki"]
    empty = sub / "empty.txt"
    empty.write_text("# nothing\n")
    assert fk.instance_list_read(str(empty)) == []
    # a path longer than any fixed line buffer stays one entry; no final newline
    deep = "/".join(["d" * 200] * 30) + "/p.fki"
    lng = sub / "long.txt"
    lng.write_text(f"{deep}\nlast.fki")
    assert fk.instance_list_read(str(lng)) == [str(sub / deep), str(sub / "last.fki")]


def test_header_whose_extents_wrap_is_rejected(fk, tmp_path):
    """number_bands = ncouls = 2^30 make 16 * nb * ncouls = 2^64, which wraps
    to 0 in 64-bit arithmetic; every other extent fits a 48 GiB sparse file,
    so only an overflow-checked extent rejects the header."""
    nb = C = 1 << 30
    path = str(tmp_path / "wrap.fki")
    size = 48 << 30
    offs = [256, 320, 384]                                  # wx, inv_igp_index, indinv
    offs.append(offs[-1] + (4 * (C + 1) + 63) // 64 * 64)  # vcoul
    offs.append(offs[-1] + 8 * C)                           # aqsmtemp (its wrapped extent would be 0)
    offs.append(offs[-1])                                   # aqsntemp
    offs.append(offs[-1])                                   # I_eps_array: 16 * ngpown * C = 16 GiB
    offs.append(offs[-1] + 16 * C)                          # wtilde_array
    assert offs[-1] + 16 * C <= size
    header = HEADER.pack(b"FKINST01", 1, 16, nb, 0, C, 1, 1, 0, 0, 0, 1e-6, 0.25, 2.5e5, 1.0, 0, 0, 0, 0,
                         *(offs + [0] * 8))
    try:
        with open(path, "wb") as f:
            f.write(header)
            f.truncate(size)
    except OSError as e:
        pytest.skip(f"no sparse file of {size} bytes here: {e}")
    with pytest.raises(fk.FkError) as e:
        fk.Instance(path)
    assert e.value.code == fk.FK_ERR_ARG
    msg = fk.last_error()
    if "mmap" in msg:
        pytest.skip(f"cannot map a {size}-byte file here: {msg}")
    assert "extent" in msg
