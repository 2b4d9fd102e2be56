"""GPU parity on AMR meshes: levels with coarse-fine faces (ghost states
interpolated from coarser leaf cells, reflux into those cells in processing
order through the coarse-target lists) and levels whose cells have refined
neighbours (zero flux) or are refined themselves (skipped)."""
import numpy as np
import pytest

import oracle
import rhydro
from amr_meshes import refined_mesh
from helpers import VARIANTS, assert_same, clone

pytestmark = pytest.mark.gpu

SPHERES = ((0.4, 0.5, 0.6, 0.3), (0.45, 0.5, 0.55, 0.15))


def _sequence(api, levels, dts):
    """Finer levels first, as the recursion of amr_step does."""
    for L in levels:
        api.set_unew(L)
    for L, dt in zip(levels, dts):
        api.godunov_fine(L, dt, 1.0 / 2 ** L)
    for L in levels:
        api.set_uold(L)


class _OracleApi:
    def __init__(self, m, p):
        self.om, self.p = oracle.OracleMesh(m), p

    def set_unew(self, L):
        self.om.set_unew(L)

    def set_uold(self, L):
        self.om.set_uold(L)

    def godunov_fine(self, L, dt, dx):
        self.om.godunov_fine(self.p, L, dt, dx)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("riemann", ["llf", "hllc"])
@pytest.mark.parametrize("mode", [rhydro.RH_DEVICE_RESIDENT, rhydro.RH_HOST_COHERENT])
def test_two_level_refinement(riemann, mode, variant):
    m0 = refined_mesh(4, SPHERES, riemann=rhydro.resolve_riemann(riemann))
    levels = [6, 5, 4]
    dts = [2e-3, 4e-3, 8e-3]
    print(f"case amr {riemann} mode={mode} octs={m0.numbl.tolist()}", flush=True)
    mg, mo = clone(m0), clone(m0)
    s = rhydro.Solver(gamma=1.4, slope_type=1, riemann=riemann, variant=variant)
    s.bind(mg, mode)
    _sequence(s, levels, dts)
    if mode == rhydro.RH_DEVICE_RESIDENT:
        s.sync_to_host()
    _sequence(_OracleApi(mo, oracle.params(1.4, 1, rhydro.resolve_riemann(riemann))), levels, dts)
    assert_same(mg.unew, mo.unew, "unew")
    assert_same(mg.uold, mo.uold, "uold")
    s.close()


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("slope_type,interpol_type", [(2, 1), (1, 2), (8, 3), (0, 0)])
def test_single_call_on_each_amr_level(slope_type, interpol_type, variant):
    m0 = refined_mesh(4, SPHERES, seed=11, interpol_type=interpol_type)
    for L in (4, 5, 6):
        print(f"case amr level {L} slope {slope_type} interpol {interpol_type}", flush=True)
        mg, mo = clone(m0), clone(m0)
        s = rhydro.Solver(gamma=1.4, slope_type=slope_type, riemann="hllc", interpol_type=interpol_type,
                          variant=variant)
        s.bind(mg)
        s.set_unew(L)
        s.godunov_fine(L, 3e-3, 1.0 / 2 ** L)
        s.sync_to_host()
        om = oracle.OracleMesh(mo)
        om.set_unew(L)
        om.godunov_fine(oracle.params(1.4, slope_type, 2, interpol_type), L, 3e-3, 1.0 / 2 ** L)
        assert_same(mg.unew, mo.unew, f"unew level {L}")
        s.close()


def test_courant_on_amr_levels_skips_refined_cells():
    m0 = refined_mesh(4, SPHERES, seed=3)
    s = rhydro.Solver(gamma=1.4)
    s.bind(m0)
    om = oracle.OracleMesh(clone(m0))
    for L in (4, 5, 6):
        print(f"case amr courant level {L}", flush=True)
        assert s.courant(L, 1.0 / 2 ** L, 1.4, 0.8) == om.courant(L, 1.0 / 2 ** L, 1.4, 0.8)
    s.close()


def test_rebind_after_refinement():
    """A topology change bumps the epoch; the tables are rebuilt."""
    m = refined_mesh(4, SPHERES[:1], extra_grids=4000, seed=5, extra_levels=1)
    s = rhydro.Solver(gamma=1.4, riemann="hllc")
    s.bind(m, rhydro.RH_HOST_COHERENT)
    s.set_unew(5)
    s.godunov_fine(5, 1e-3, 1 / 32)
    from amr_meshes import cell_centres

    cells, cen = cell_centres(m, 5)
    inside = ((cen - np.array([0.4, 0.5, 0.6])) ** 2).sum(1) < 0.1 ** 2
    oracle.refine(m, oracle.params(1.4, 1, 2), 5, cells[inside & (m.son[cells - 1] == 0)])
    mo = clone(m)
    s.bind(m, rhydro.RH_HOST_COHERENT)
    print("case amr rebind", flush=True)
    for api in (s, _OracleApi(mo, oracle.params(1.4, 1, 2))):
        api.set_unew(6)
        api.set_unew(5)
        api.godunov_fine(6, 5e-4, 1 / 64)
        api.godunov_fine(5, 1e-3, 1 / 32)
    assert_same(m.unew, mo.unew, "unew after rebind")
    s.close()


@pytest.mark.parametrize("variant", VARIANTS)
def test_checkerboard_refinement_many_reflux_targets(variant):
    """Every other cell of a 32^3 level refined in a 3D checkerboard: each
    fine oct is surrounded by coarser leaf cells on all six faces, and the
    16384 unrefined cells are all reflux targets with 24 contributions each
    -- enough targets for the eight-lanes-per-target reflux form (the smaller
    meshes above use the wave-per-target form)."""
    from amr_meshes import cell_centres

    base = 6
    ng = int(rhydro.lib().rh_uniform_ngrid(1, 1, 1, base))
    m0 = rhydro.uniform_mesh(base, ngridmax=ng + 16384, nlevelmax=base + 1)
    rhydro.init_lognormal(m0, base, 13, 1.4, 0.3, 0.3)
    cells, cen = cell_centres(m0, base)
    ijk = np.floor(cen * 2 ** (base - 1)).astype(np.int64)
    chosen = cells[(ijk.sum(1) % 2) == 0]
    assert chosen.size == 16384
    oracle.refine(m0, oracle.params(1.4, 1, 2), base, chosen)
    m0.unew[:] = m0.uold
    levels = [base + 1, base]
    dts = [1e-3, 2e-3]
    print(f"case amr checkerboard variant={variant} octs={m0.numbl.tolist()}", flush=True)
    mg, mo = clone(m0), clone(m0)
    s = rhydro.Solver(gamma=1.4, slope_type=1, riemann="hllc", variant=variant)
    s.bind(mg)
    _sequence(s, levels, dts)
    s.sync_to_host()
    _sequence(_OracleApi(mo, oracle.params(1.4, 1, 2)), levels, dts)
    assert_same(mg.unew, mo.unew, "unew")
    assert_same(mg.uold, mo.uold, "uold")
    s.close()


def test_autotune_on_refined_levels_leaves_state_and_results():
    """rh_autotune on levels with coarser neighbours and refined cells (its
    trials without the Courant epilogue, pre-update state from uold) leaves
    uold and unew as they were, and the tuned context still equals the
    oracle value for value."""
    m0 = refined_mesh(4, SPHERES, riemann=rhydro.resolve_riemann("hllc"))
    levels = [6, 5, 4]
    dts = [2e-3, 4e-3, 8e-3]
    mg, mo = clone(m0), clone(m0)
    s = rhydro.Solver(gamma=1.4, slope_type=1, riemann="hllc")
    s.bind(mg, rhydro.RH_DEVICE_RESIDENT)
    for L in (6, 5):
        print(f"case autotune amr level {L} octs={m0.numbl.tolist()}", flush=True)
        best, ms = s.autotune(L, 2e-3, 1.0 / 2 ** L, reps=1)
        assert best in range(rhydro.RH_KERNEL_VARIANTS) and s.variant() == best
        assert (ms > 0).all()
        s.sync_to_host()
        assert_same(mg.uold, m0.uold, "uold after autotune")
        assert_same(mg.unew, m0.unew, "unew after autotune")
    _sequence(s, levels, dts)
    s.sync_to_host()
    _sequence(_OracleApi(mo, oracle.params(1.4, 1, rhydro.resolve_riemann("hllc"))), levels, dts)
    assert_same(mg.unew, mo.unew, "unew")
    assert_same(mg.uold, mo.uold, "uold")
    s.close()
