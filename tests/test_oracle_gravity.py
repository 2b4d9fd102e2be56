"""Pins of the oracle's gravity source terms (HydroSolver.cpp:980-1043) by
closed forms on dyadic states, where every operation is exact."""
import numpy as np

import oracle
import rhydro
from amr_meshes import refined_mesh


def _uniform(level, rho, mom, E):
    m = rhydro.uniform_mesh(level)
    u = m.state("uold")
    u[0], u[1], u[2], u[3], u[4] = rho, mom[0], mom[1], mom[2], E
    m.unew[:] = m.uold
    return m


def test_kick_into_unew_closed_form():
    m = _uniform(3, 2.0, (0.0, 0.0, 0.0), 3.0)
    f = np.zeros(3 * m.ncell)
    f.reshape(3, -1)[:] = np.array([0.5, -0.25, 1.0])[:, None]
    om = oracle.OracleMesh(m, f=f)
    om.add_gravity_source_terms(3, 0.125)
    cells = m.cells_of_level(3)
    u = m.state("unew")[:, cells]
    np.testing.assert_array_equal(u[1], 2.0 * 0.5 * 0.125)
    np.testing.assert_array_equal(u[2], 2.0 * -0.25 * 0.125)
    np.testing.assert_array_equal(u[3], 2.0 * 1.0 * 0.125)
    # E gains rho f^2 dt^2 / 2 per dimension (uold momentum is zero)
    np.testing.assert_array_equal(u[4], 3.0 + 0.5 * 2.0 * 0.125 ** 2 * (0.25 + 0.0625 + 1.0))
    np.testing.assert_array_equal(m.state("unew")[0, cells], 2.0)


def test_kick_uses_uold_momentum_for_the_work_term():
    m = _uniform(3, 1.0, (0.5, 0.0, 0.0), 2.0)
    m.state("unew")[1] = 0.25  # unew differs from uold: the work term reads uold
    f = np.zeros(3 * m.ncell)
    f.reshape(3, -1)[0] = 1.0
    oracle.OracleMesh(m, f=f).add_gravity_source_terms(3, 0.5)
    cells = m.cells_of_level(3)
    np.testing.assert_array_equal(m.state("unew")[1, cells], 0.25 + 0.5)
    np.testing.assert_array_equal(m.state("unew")[4, cells], 2.0 + 0.5 * 1.0 * 0.5 + 0.5 * 1.0 * 0.25)


def test_momentum_clamp_and_zero_field():
    m = _uniform(3, 1.0, (30.0, -30.0, 1.0), 500.0)
    om = oracle.OracleMesh(m)  # no field: f = 0, but the clamp still applies
    om.add_gravity_source_terms(3, 0.1)
    cells = m.cells_of_level(3)
    np.testing.assert_array_equal(m.state("unew")[1:4, cells].T, np.array([10.0, -10.0, 1.0])[None, :].repeat(len(cells), 0))
    np.testing.assert_array_equal(m.state("unew")[4, cells], 500.0)


def test_synchro_keeps_internal_energy():
    m = _uniform(3, 1.0, (0.5, 0.0, 0.0), 1.0)
    f = np.zeros(3 * m.ncell)
    f.reshape(3, -1)[0] = 1.0
    oracle.OracleMesh(m, f=f).synchro_hydro_fine(3, 0.25, 1.4)
    cells = m.cells_of_level(3)
    u = m.state("uold")[:, cells]
    np.testing.assert_array_equal(u[1], 0.75)
    np.testing.assert_array_equal(u[4], (1.0 - 0.5 * 0.25) + 0.5 * 0.5625)
    np.testing.assert_array_equal(u[0], 1.0)


def test_refined_cells_are_skipped():
    m = refined_mesh(4, ((0.5, 0.5, 0.5, 0.2),), seed=3)
    before_new, before_old = m.unew.copy(), m.uold.copy()
    f = np.random.default_rng(1).normal(0, 1, 3 * m.ncell)
    om = oracle.OracleMesh(m, f=f)
    om.add_gravity_source_terms(4, 0.01)
    om.synchro_hydro_fine(4, 0.01, 1.4)
    cells = m.cells_of_level(4)
    refined = cells[m.son[cells] > 0]
    leaf = cells[m.son[cells] == 0]
    assert len(refined) > 0 and len(leaf) > 0
    np.testing.assert_array_equal(m.state("unew")[:, refined], before_new.reshape(5, -1)[:, refined])
    np.testing.assert_array_equal(m.state("uold")[:, refined], before_old.reshape(5, -1)[:, refined])
    assert (m.state("uold")[1, leaf] != before_old.reshape(5, -1)[1, leaf]).any()
