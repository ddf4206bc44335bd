"""GPU parity of batched FF (ff_batch_*): every problem of a batch of uneven
problems gets bit for bit what ff_kernel gives it alone with the same
configuration, and the oracle's result to 1e-12 (FP64) or 5e-6 (split
BF16); graph replays see changed inputs with eager work between them."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(40, 300, 70, 3), (7, 64, 64, 5), (129, 500, 33, 2), (1, 17, 5, 1), (64, 1024, 256, 4)]


def _torch():
    import torch
    assert torch.cuda.is_available(), "no ROCm device visible"
    return torch


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _problems(fko, seed0=0):
    torch = _torch()
    d = torch.device("cuda:0")
    f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(d).to(dt)  # noqa: E731
    ps, dev = [], []
    for i, (nb, C, G, nf) in enumerate(SHAPES):
        p = fko.ff_random(nb, C, G, nf, seed=seed0 + i)
        if i == 2:                                 # permuted and skipped columns
            inv = p.inv_igp_index.copy()
            inv[[0, 4]] = inv[[4, 0]]
            ind = p.indinv.copy()
            ind[inv[6]] = -1
            p.inv_igp_index, p.indinv = inv, ind
        ps.append(p)
        dev.append(dict(nb=nb, ncouls=C, ngpown=G, nfreq=nf, aqsm=f(p.aqsm, torch.complex128),
                        aqsn=f(p.aqsn, torch.complex128), I_eps=f(p.I_eps, torch.complex128),
                        vcoul=f(p.vcoul, torch.float64), inv_igp_index=p.inv_igp_index, indinv=p.indinv,
                        achs_freq=i % nf))
    return ps, dev


def _single(fk, q, precision, config):
    out = fk.ff_kernel(q["nb"], q["ncouls"], q["ngpown"], q["nfreq"], q["aqsm"], q["aqsn"], q["I_eps"], q["vcoul"],
                       q["inv_igp_index"], q["indinv"], achs_freq=q["achs_freq"], precision=precision,
                       config=config)
    return out[0].cpu().numpy(), out[1].cpu().numpy(), complex(out[2].cpu()[0])


@pytest.mark.parametrize("precision,config", [("fp64", 1), ("fp64", 2), ("fp64", 3), ("mixed", 1), ("mixed", 3)])
def test_batch_equals_single_calls_and_oracle(fk, fko, precision, config):
    torch = _torch()
    print(f"ff batch {precision} config {config}", flush=True)
    ps, dev = _problems(fko)
    b = fk.FfBatch(dev, precision=precision, config=config)
    assert b.config == config and b.workspace_bytes > 0
    t, v, a, ms = b.run(time_it=True)
    torch.cuda.synchronize()
    assert ms > 0
    tol = 1e-12 if precision == "fp64" else 5e-6
    for i, (p, q) in enumerate(zip(ps, dev)):
        st, sv, sa = _single(fk, q, precision, config)
        assert np.array_equal(t[i].cpu().numpy(), st), i
        assert np.array_equal(v[i].cpu().numpy(), sv), i
        assert complex(a[i].cpu()[0]) == sa, i
        rt, rv, ra = fko.ff(p, achs_freq=q["achs_freq"])
        assert _rel(st, rt) <= tol and _rel(sv, rv) <= tol, (i, _rel(st, rt))
        assert abs(sa - ra) <= tol * abs(ra), i


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotuned_batch(fk, fko, precision):
    torch = _torch()
    ps, dev = _problems(fko, seed0=20)
    b = fk.FfBatch(dev, precision=precision, config=0)
    ncfg = 3
    assert 1 <= b.config <= ncfg
    t, _, _, _ = b.run()
    torch.cuda.synchronize()
    for i, q in enumerate(dev):
        st, _, _ = _single(fk, q, precision, b.config)
        assert np.array_equal(t[i].cpu().numpy(), st), i
    print(f"ff batch {precision}: autotuned config {b.config}", flush=True)


def test_replays_see_new_inputs_and_null_outputs(fk, fko):
    torch = _torch()
    ps, dev = _problems(fko, seed0=40)
    b = fk.FfBatch(dev, precision="fp64", config=2)
    for rep in range(3):
        # eager work between replays, and new values in the same arrays
        dev[1]["I_eps"].mul_(1.5 if rep else 1.0)
        dev[4]["aqsn"].add_(0.125 * rep)
        torch.cuda.synchronize()
        t, v, a, _ = b.run(want_schDv=(rep != 1), want_achs=(rep != 2))
        torch.cuda.synchronize()
        assert (v is None) == (rep == 1) and (a is None) == (rep == 2)
        for i, q in enumerate(dev):
            st, sv, sa = _single(fk, q, "fp64", 2)
            assert np.array_equal(t[i].cpu().numpy(), st), (rep, i)
            if v is not None:
                assert np.array_equal(v[i].cpu().numpy(), sv), (rep, i)
            if a is not None:
                assert complex(a[i].cpu()[0]) == sa, (rep, i)
    b.close()
