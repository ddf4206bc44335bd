"""Pin the oracle's gppsum = 1 band loop (oracle/fko_gpp_seq.c) before it
checks the device: golden vectors written by the reference's own compiled
gppKerSeq.cpp functions (flagOCC_solver / noflagOCC_solver, :76-261, called
as its main calls them) and program (its printed achtemp, gppsum = 1), and
the block bounds of its unoccupied loop against a literal walk of the
reference's loop structure (:438-450 with :209-241)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _problem(fko, d):
    return fko.GppProblem(int(d["nb"]), int(d["nv"]), int(d["ncouls"]), int(d["ngpown"]), d["wx"],
                          d["aqsm"], d["aqsn"], d["I_eps"], d["wtilde"], d["vcoul"],
                          d["inv_igp_index"], d["indinv"], occ=float(d["occ"]))


def _close(got, ref, rtol, what):
    scale = np.max(np.abs(ref))
    for g, r in zip(got, ref):
        assert abs(g - r) <= rtol * abs(r) + 1e-2 * rtol * scale, (what, g, r)


@pytest.mark.parametrize("case", ["random", "blocks", "break"])
def test_gppsum_oracle_matches_reference_functions(fko, case):
    d = np.load(os.path.join(GOLDEN, f"gppsum_{case}.npz"))
    ach, asx, n1 = fko.gpp_gppsum(_problem(fko, d), nthreads=2, want_n1loc=True)
    _close(ach, d["achtemp"], 1e-13, f"{case} achtemp")
    _close(asx, d["asxtemp"], 1e-13, f"{case} asxtemp")
    _close(n1, d["acht_n1_loc"], 1e-13, f"{case} acht_n1_loc")


def test_gppsum_oracle_matches_reference_program(fko):
    """gppKerSeq <nb> <nv> <ncouls> <npg> 1 with its constant inputs and
    its column map (i+1)*ncouls/ngpown, printed at 15 digits."""
    d = np.load(os.path.join(GOLDEN, "gppsum_main.npz"))
    for (nb, nv, C, npg), want in zip(d["shapes"], d["achtemp"]):
        p = fko.gpp_constant(int(nb), int(nv), int(C), int(npg), map_kind="cpp")
        ach, _, _ = fko.gpp_gppsum(p)
        _close(ach, want, 1e-13, f"program {nb}/{nv}/{C}/{npg}")


def _reference_block_walk(igp):
    """{ig: kind} of the reference's unoccupied loop for one column, walked
    as its main and noflagOCC_solver walk it: blocks igbeg = 0, 512, ...
    < igmax, igend = min(igbeg + 511, igmax), ig in [igbeg, min(igend,
    igmax - 1)), then ig = igmax (m1 term only) if igend == igmax - 1."""
    igmax, out = igp, {}
    igbeg = 0
    while igbeg < igmax:
        igend = min(igbeg + 512 - 1, igmax)
        for ig in range(igbeg, min(igend, igmax - 1)):
            out[ig] = 1
        if igend == igmax - 1:
            out[igmax] = 2
        igbeg += 512
    return out


def test_gppsum_block_bounds_against_the_reference_walk(fko):
    for igp in list(range(0, 1600)) + [2047, 2048, 2049, 4095, 4096, 32767]:
        walk = _reference_block_walk(igp)
        for ig in range(0, igp + 2):
            assert fko.gppsum_unocc_kind(ig, igp) == walk.get(ig, 0), (ig, igp)


def test_gppsum_fixtures_exercise_each_rule(fko):
    """The block rules matter to the fixture's points: with every band
    unoccupied, moving the near-pole wtilde of one decided point changes
    the oracle's achtemp exactly when the rule keeps that point."""
    d = np.load(os.path.join(GOLDEN, "gppsum_blocks.npz"))
    p = _problem(fko, d)
    p.nv = 0
    base, _, _ = fko.gpp_gppsum(p)
    igp = fko.resolve_igp(p.ncouls, p.ngpown, p.inv_igp_index, p.indinv)
    kinds = set()
    for m, ig in ((4, 512), (8, 1024), (4, 511), (8, 1023), (5, 511), (9, 1022), (7, 510)):
        q = _problem(fko, d)
        q.nv = 0
        w = q.wtilde.copy()
        w[m, ig] = 3.15 + 0.03j
        q.wtilde = w
        moved, _, _ = fko.gpp_gppsum(q)
        kind = fko.gppsum_unocc_kind(ig, int(igp[m]))
        kinds.add(kind)
        assert (not np.array_equal(moved, base)) == (kind != 0), (m, ig, int(igp[m]))
    assert kinds == {0, 1, 2}
