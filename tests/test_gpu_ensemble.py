"""The batched (ensemble) form through the C-ABI: every box of an ensemble
must come out exactly as it does alone -- bit-identical to the oracle in the
FP64 form, and to its own single-mesh run in the mixed form -- with its own
time steps."""
import numpy as np
import pytest

import oracle
import rhydro
from helpers import VARIANTS, assert_same, clone, oracle_params

pytestmark = pytest.mark.gpu

# (nx, level): sides 4, 6, 8, 10, 12, 14, 16 cells; dx a power of two only
# for nx = 1, so both the exact-reciprocal and the dividing FP64 kernels run.
SHAPES = [(1, 3), (3, 2), (1, 4), (5, 2), (3, 3), (7, 2), (1, 5)]


def _boxes(seed, shapes=SHAPES, gamma=1.4):
    rng = np.random.default_rng(seed)
    out = []
    for k, (nx, L) in enumerate(shapes):
        m = rhydro.uniform_mesh(L, nx=nx, ny=nx, nz=nx)
        if k % 2:
            rhydro.init_lognormal(m, L, seed + k, gamma, 0.4, 0.4)
        else:
            n = rng.normal(size=3)
            n /= np.linalg.norm(n)
            d = rng.uniform(0.1, 1.0, 4)
            rhydro.init_shock_tube(m, n, [0.5, 0.5, 0.5], [d[0], 0, 0, 0, d[1]], [d[2], 0, 0, 0, d[3]], gamma)
        out.append((m, L, 1.0 / (nx * 2 ** L)))
    return out


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("riemann", ["llf", "hll", "hllc", "acoustic"])
def test_ensemble_fp64_matches_oracle_per_box(riemann, variant):
    boxes = _boxes(100)
    meshes = [clone(m) for m, _, _ in boxes]
    levels = [L for _, L, _ in boxes]
    dxs = [dx for _, _, dx in boxes]
    print(f"case ensemble fp64 {riemann} variant {variant} boxes {len(boxes)}", flush=True)
    s = rhydro.Solver(gamma=1.4, slope_type=1, riemann=riemann, variant=variant)
    s.bind_ensemble(meshes)
    dts = s.run_ensemble(levels, dxs, 0.8, 3)
    s.sync_to_host()
    p = oracle_params(s)
    s.close()
    for b, (m0, L, dx) in enumerate(boxes):
        mo = clone(m0)
        dto = oracle.OracleMesh(mo).run_uniform(p, L, dx, 0.8, 3)
        assert_same(dts[:, b], dto, f"box {b} dt")
        assert_same(meshes[b].uold, mo.uold, f"box {b} uold")
        assert_same(meshes[b].unew, mo.unew, f"box {b} unew")


@pytest.mark.parametrize("riemann", ["hll", "hllc"])
def test_ensemble_mixed_matches_single_box_runs(riemann):
    boxes = _boxes(200)
    meshes = [clone(m) for m, _, _ in boxes]
    print(f"case ensemble mixed {riemann}", flush=True)
    s = rhydro.Solver(gamma=1.4, slope_type=2, riemann=riemann, precision=rhydro.RH_PREC_MIXED)
    s.bind_ensemble(meshes)
    dts = s.run_ensemble([L for _, L, _ in boxes], [dx for _, _, dx in boxes], 0.8, 4)
    s.sync_to_host()
    for b, (m0, L, dx) in enumerate(boxes):
        m1 = clone(m0)
        s.bind(m1)
        dt1 = s.run_uniform(L, dx, 0.8, 4)
        s.sync_to_host()
        assert_same(dts[:, b], dt1, f"box {b} dt")
        assert_same(meshes[b].uold, m1.uold, f"box {b} uold")
    s.close()


def test_ensemble_host_coherent_and_repeated_runs():
    boxes = _boxes(300, shapes=[(1, 3), (3, 2), (5, 2)])
    meshes = [clone(m) for m, _, _ in boxes]
    levels = [L for _, L, _ in boxes]
    dxs = [dx for _, _, dx in boxes]
    print("case ensemble host-coherent", flush=True)
    s = rhydro.Solver(gamma=1.4, riemann="hllc")
    s.bind_ensemble(meshes, rhydro.RH_HOST_COHERENT)
    d1 = s.run_ensemble(levels, dxs, 0.8, 2)
    # The host may edit a box between calls in host-coherent mode.
    meshes[1].uold[:] = boxes[1][0].uold
    meshes[1].unew[:] = boxes[1][0].unew
    d2 = s.run_ensemble(levels, dxs, 0.8, 2)
    s.close()
    p = oracle.params(1.4, 1, 2)
    for b, (m0, L, dx) in enumerate(boxes):
        mo = clone(m0)
        om = oracle.OracleMesh(mo)
        first = om.run_uniform(p, L, dx, 0.8, 2)
        if b == 1:  # restarted from its initial state
            mo = clone(m0)
            om = oracle.OracleMesh(mo)
        second = om.run_uniform(p, L, dx, 0.8, 2)
        assert_same(d1[:, b], first, f"box {b} dt, first call")
        assert_same(d2[:, b], second, f"box {b} dt, second call")
        assert_same(meshes[b].uold, mo.uold, f"box {b} uold")


def test_ensemble_leaf_cells_and_errors():
    boxes = _boxes(400, shapes=[(1, 3), (3, 2)])
    meshes = [clone(m) for m, _, _ in boxes]
    s = rhydro.Solver(gamma=1.4)
    s.bind_ensemble(meshes)
    print("case ensemble errors", flush=True)
    assert s.ensemble_leaf_cells([3, 2]) == 4 ** 3 + 6 ** 3
    with pytest.raises(rhydro.RhError):
        s.godunov_fine(3, 1e-3, 1 / 8)  # single-mesh entry points refuse an ensemble
    with pytest.raises(rhydro.RhError):
        s.run_ensemble([3, 9], [1 / 8, 1 / 12], 0.8, 1)  # level beyond the box's nlevelmax
    with pytest.raises(rhydro.RhError):
        s.run_ensemble([3, 2], [1 / 8, 0.0], 0.8, 1)
    # Binding a single mesh afterwards replaces the ensemble.
    m = clone(boxes[0][0])
    s.bind(m)
    s.run_uniform(3, 1 / 8, 0.8, 1)
    with pytest.raises(rhydro.RhError):
        s.run_ensemble([3, 2], [1 / 8, 1 / 12], 0.8, 1)
    s.close()


def test_single_mesh_non_power_of_two_dx():
    """nx = 3: dx = 1/(3*2^L) is not a power of two, so the FP64 kernels
    divide; with nx = 1 they multiply by the exact reciprocal.  Both match
    the oracle bit for bit."""
    for nx, L in ((3, 3), (1, 4)):
        m0 = rhydro.uniform_mesh(L, nx=nx, ny=nx, nz=nx)
        rhydro.init_lognormal(m0, L, 17, 1.4, 0.5, 0.5)
        mg, mo = clone(m0), clone(m0)
        dx = 1.0 / (nx * 2 ** L)
        print(f"case single nx={nx} level={L}", flush=True)
        s = rhydro.Solver(gamma=1.4, slope_type=2, riemann="hllc")
        s.bind(mg)
        dtg = s.run_uniform(L, dx, 0.8, 3)
        s.sync_to_host()
        dto = oracle.OracleMesh(mo).run_uniform(oracle_params(s), L, dx, 0.8, 3)
        s.close()
        assert_same(dtg, dto, "dt")
        assert_same(mg.uold, mo.uold, "uold")


def test_benchmark_ensemble_96_boxes_fp64_step_per_box():
    """The benchmark's batched workload: 96 independent periodic boxes of
    16^3-64^3 cells, shock tubes and Sedov blasts (seed 20261202), advanced
    together by one FP64 step, each box with its own time step.  Every box
    equals the oracle run on it alone, time step included."""
    boxes = rhydro.box_ensemble(96, seed=20261202)
    assert {b.kind for b in boxes} == {"sod", "sedov"}
    meshes = [clone(b.mesh) for b in boxes]
    print("case ensemble 96 boxes fp64", flush=True)
    s = rhydro.Solver(gamma=1.4, slope_type=1, riemann="hllc")
    s.bind_ensemble(meshes)
    dts = s.run_ensemble([b.level for b in boxes], [b.dx for b in boxes], 0.8, 1)
    s.sync_to_host()
    p = oracle_params(s)
    s.close()
    for k, b in enumerate(boxes):
        mo = clone(b.mesh)
        dto = oracle.OracleMesh(mo).run_uniform(p, b.level, b.dx, 0.8, 1)
        assert_same(dts[:, k], dto, f"box {k} ({b.kind}) dt")
        assert_same(meshes[k].uold, mo.uold, f"box {k} ({b.kind}) uold")
        assert_same(meshes[k].unew, mo.unew, f"box {k} ({b.kind}) unew")


def test_benchmark_ensemble_96_boxes_fp64_eleven_steps():
    """The benchmark's ensemble over warmup + timed steps (two calls, 1 + 10
    steps): the batched FP64 run equals the oracle on some of its boxes -- Sedov blasts
    whose state the reference scheme blows up within these steps (negative
    pressures, speeds of 1e10), Sedov blasts that stay regular, and shock
    tubes -- time steps included."""
    boxes = rhydro.box_ensemble(96, seed=20261202)
    meshes = [clone(b.mesh) for b in boxes]
    print("case ensemble 96 boxes fp64, 11 steps", flush=True)
    s = rhydro.Solver(gamma=1.4, slope_type=1, riemann="hllc")
    s.bind_ensemble(meshes)
    lv, dx = [b.level for b in boxes], [b.dx for b in boxes]
    # as the benchmark calls it: one warmup step, then ten
    dts = np.concatenate([s.run_ensemble(lv, dx, 0.8, 1), s.run_ensemble(lv, dx, 0.8, 10)])
    s.sync_to_host()
    p = oracle_params(s)
    s.close()
    for k in (29, 49, 55, 0, 1, 86):
        b = boxes[k]
        mo = clone(b.mesh)
        print(f"box {k} ({b.kind}, side {b.nx * 2 ** (b.level - 1)})", flush=True)
        dto = oracle.OracleMesh(mo).run_uniform(p, b.level, b.dx, 0.8, 11)
        assert_same(dts[:, k], dto, f"box {k} ({b.kind}) dt")
        assert_same(meshes[k].uold, mo.uold, f"box {k} ({b.kind}) uold")
