"""CPU checks of the AMR test-mesh generator (the oracle's make_grid_fine
restatement): level lists stay inside nlevelmax, octs are consistent."""
import numpy as np
import pytest

import oracle
from amr_meshes import cell_centres, refined_mesh


def test_refine_beyond_nlevelmax_is_rejected():
    m = refined_mesh(4, ((0.4, 0.5, 0.6, 0.3),), extra_grids=2000, seed=5)
    cells, _ = cell_centres(m, 5)
    with pytest.raises(ValueError):
        oracle.refine(m, oracle.params(1.4, 1, 2), 5, cells[:8])


def test_refined_mesh_structure():
    m = refined_mesh(4, ((0.4, 0.5, 0.6, 0.3), (0.45, 0.5, 0.55, 0.15)), extra_levels=1)
    assert m.nlevelmax == 7 and len(m.headl) == 8
    for L in (5, 6):
        octs = m.level_octs(L)
        assert len(octs) == m.numbl[L] > 0
        for g in octs:
            fc = m.father[g - 1]
            assert m.son[fc - 1] == g  # father/son agree
    # Note: a state at rest does not stay at rest on an AMR mesh in the
    # reference: the coarse face next to a refined cell carries zero flux
    # (HydroSolver.cpp:125-128) and the fine level refluxes 4 x 1/8 of its
    # face fluxes (:476-482), so momentum appears at coarse-fine interfaces.
    m.uold.reshape(5, -1)[:] = np.array([1.0, 0, 0, 0, 2.5])[:, None]
    m.unew[:] = m.uold
    om = oracle.OracleMesh(m)
    p = oracle.params(1.4, 1, 2)
    for L in (6, 5, 4):
        om.set_unew(L)
    for L in (6, 5, 4):
        om.godunov_fine(p, L, 1e-3, 1.0 / 2 ** L)
    cells = np.concatenate([m.cells_of_level(L, leaf_only=True) for L in (4, 5, 6)])
    assert np.isfinite(m.state("unew")[:, cells]).all()
    assert np.abs(m.state("unew")[1:4, cells]).max() > 0
