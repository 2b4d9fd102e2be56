"""Pins of the oracle's restatement of the adaptive run (tree updater, level
and free lists, subcycled amr_step)."""
import numpy as np
import pytest

import oracle
import rhydro


def _lists(s):
    """Grids of every level list, in list order."""
    out = {}
    for L in range(1, s["nlevelmax"] + 1):
        g, seq = int(s["headl"][L]), []
        while g > 0:
            seq.append(g)
            g = int(s["next"][g - 1])
        out[L] = seq
    return out


def _depth(s, g):
    d, nc, ng = 0, s["ncoarse"], s["ngridmax"]
    while True:
        f = int(s["father"][g - 1])
        d += 1
        if f <= nc:
            return d
        g = (f - nc - 1) % ng + 1


def test_uniform_run_equals_the_uniform_driver():
    """levelmin == levelmax: no tree changes, and each coarse step is one
    courant -> set_unew -> godunov_fine -> set_uold of the level, so the run
    must equal the oracle's uniform stepping on the reference's uniform mesh."""
    L = 4
    regs = rhydro.sedov_regions(L)
    a = oracle.OracleAmr(levelmin=L, levelmax=L, ngridmax=rhydro.lib().rh_uniform_ngrid(1, 1, 1, L), regions=regs)
    m = rhydro.uniform_mesh(L)
    rhydro.init_regions(m, regs, 1.4)
    s0 = a.state()
    np.testing.assert_array_equal(s0["uold"], m.uold)
    np.testing.assert_array_equal(s0["son"], m.son)
    np.testing.assert_array_equal(s0["nbor"], m.nbor)
    dts = oracle.OracleMesh(m).run_uniform(oracle.params(1.4, 1, 2), L, 1.0 / 2 ** L, 0.8, 3)
    for k in range(3):
        assert a.step(1) == 0
        assert a.state()["dtnew"][L] == dts[k]
    s = a.state()
    np.testing.assert_array_equal(s["uold"], m.uold)
    assert s["nstep"] == 3 and s["t"] == sum(dts)
    assert s["cell_updates"] == 3 * 8 ** (L - 1)


@pytest.fixture(scope="module")
def sedov_probe():
    """The reference's AMR Sedov probe: levelmin 6, levelmax 9,
    err_grad_d = err_grad_p = 0.1, the 8-cell blast at level 6."""
    a = oracle.OracleAmr(levelmin=6, levelmax=9, ngridmax=60000, err_grad_d=0.1, err_grad_p=0.1,
                         regions=rhydro.sedov_regions(6))
    init = a.state()
    assert a.step(2) == 0
    return init, a.state()


def test_probe_level_counts(sedov_probe):
    init, s = sedov_probe
    assert init["numbl"][6] == 32 ** 3 // 8
    # The reference run logged 160 octs in list 7 after two coarse steps.
    assert s["numbl"][7] == 160
    assert s["numbl"][8] > 0 and s["numbl"][9] > 0


def test_tree_invariants(sedov_probe):
    _, s = sedov_probe
    lists = _lists(s)
    seen = set()
    for L, seq in lists.items():
        assert len(seq) == s["numbl"][L]
        assert not (set(seq) & seen)
        seen |= set(seq)
        for g in seq[:200]:
            assert _depth(s, g) == L - 1  # list L holds octs at father-chain depth L-1
    nc, ng = s["ncoarse"], s["ngridmax"]
    for L, seq in lists.items():
        for g in seq:
            f = int(s["father"][g - 1])
            assert s["son"][f - 1] == g
    # Every son pointer belongs to a listed grid.
    sons = s["son"][s["son"] > 0]
    assert set(sons.tolist()) <= seen


def test_refined_region_surrounds_the_blast(sedov_probe):
    _, s = sedov_probe
    lists = _lists(s)
    xg = s["xg"].reshape(3, -1)
    for L in (7, 8, 9):
        centres = np.array([xg[:, g - 1] / 1.0 for g in lists[L]])
        # xg is in coarse-cell units (nx = 1): the refined octs cluster around
        # the box centre where the energy was deposited.
        r = np.sqrt(((centres - 0.5) ** 2).sum(1))
        assert r.max() < 0.35, (L, r.max())
