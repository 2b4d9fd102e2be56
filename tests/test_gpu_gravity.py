"""GPU parity of the gravity source terms (add_gravity_source_terms,
synchro_hydro_fine) against the oracle, value-exact."""
import numpy as np
import pytest

import oracle
import rhydro
from amr_meshes import refined_mesh
from helpers import assert_same, clone

pytestmark = pytest.mark.gpu


def _field(m, seed, scale=3.0):
    return np.random.default_rng(seed).normal(0, scale, 3 * m.ncell)


@pytest.mark.parametrize("mode", [rhydro.RH_DEVICE_RESIDENT, rhydro.RH_HOST_COHERENT])
@pytest.mark.parametrize("with_field", [True, False])
def test_gravity_uniform(mode, with_field):
    level = 5
    m0 = rhydro.uniform_mesh(level)
    rhydro.init_lognormal(m0, level, 21, 1.4, 0.5, 2.0)  # large velocities: the clamps act
    f = _field(m0, 5, 30.0) if with_field else None
    mg, mo = clone(m0), clone(m0)
    print(f"case gravity uniform mode={mode} field={with_field}", flush=True)
    s = rhydro.Solver(gamma=1.4)
    s.bind(mg, mode, f=f)
    om = oracle.OracleMesh(mo, f=f)
    for dt in (1e-3, 0.05):
        s.add_gravity_source_terms(level, dt)
        om.add_gravity_source_terms(level, dt)
        s.synchro_hydro_fine(level, dt)
        om.synchro_hydro_fine(level, dt, 1.4)
    if mode == rhydro.RH_DEVICE_RESIDENT:
        s.sync_to_host()
    assert_same(mg.unew, mo.unew, "unew")
    assert_same(mg.uold, mo.uold, "uold")
    s.close()


def test_gravity_amr_levels():
    m0 = refined_mesh(4, ((0.4, 0.5, 0.6, 0.3), (0.45, 0.5, 0.55, 0.15)), seed=9)
    f = _field(m0, 6)
    for L in (4, 5, 6):
        print(f"case gravity amr level {L}", flush=True)
        mg, mo = clone(m0), clone(m0)
        s = rhydro.Solver(gamma=1.4)
        s.bind(mg, f=f)
        s.add_gravity_source_terms(L, 0.02)
        s.synchro_hydro_fine(L, 0.02)
        s.sync_to_host()
        s.close()
        om = oracle.OracleMesh(mo, f=f)
        om.add_gravity_source_terms(L, 0.02)
        om.synchro_hydro_fine(L, 0.02, 1.4)
        assert_same(mg.unew, mo.unew, f"unew level {L}")
        assert_same(mg.uold, mo.uold, f"uold level {L}")


def test_host_coherent_refreshes_the_field():
    """In host-coherent mode the host may change f between calls (a gravity
    solver runs on the host); every call uses the current values."""
    level = 4
    m0 = rhydro.uniform_mesh(level)
    rhydro.init_lognormal(m0, level, 8, 1.4)
    f = _field(m0, 1)
    mg, mo = clone(m0), clone(m0)
    s = rhydro.Solver(gamma=1.4)
    s.bind(mg, rhydro.RH_HOST_COHERENT, f=f)
    om = oracle.OracleMesh(mo, f=f)
    print("case gravity field refresh", flush=True)
    s.add_gravity_source_terms(level, 0.01)
    om.add_gravity_source_terms(level, 0.01)
    f[:] = _field(m0, 2)  # same array, new values
    s.add_gravity_source_terms(level, 0.01)
    om.add_gravity_source_terms(level, 0.01)
    assert s.courant(level, 1 / 16, 1.4, 0.8) == om.courant(level, 1 / 16, 1.4, 0.8)
    assert_same(mg.unew, mo.unew, "unew")
    s.close()
