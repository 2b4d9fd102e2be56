"""Multi-step runs replayed from captured HIP graphs (rh_set_graphs): every
replay must give what the eager launches give, bit for bit, with eager work
and stream synchronisations of another context between the replays; a
change of parameters captures a new graph; the timers switch graphs off.
The eager runs are themselves checked against the oracle."""
import numpy as np
import pytest

import oracle
import rhydro
from helpers import assert_same, clone, oracle_params

pytestmark = pytest.mark.gpu

LEVEL = 5
DX = 1.0 / 2**LEVEL


@pytest.fixture(scope="module")
def lognormal5():
    m = rhydro.uniform_mesh(LEVEL)
    rhydro.init_lognormal(m, LEVEL, 77, 1.4, 0.5, 0.5)
    return m


def _eager_uniform(m0, cf, nsteps, precision):
    m = clone(m0)
    s = rhydro.Solver(gamma=1.4, slope_type=1, riemann="hllc", precision=precision)
    s.graphs(False)
    s.bind(m, rhydro.RH_HOST_COHERENT)
    dts = s.run_uniform(LEVEL, DX, cf, nsteps)
    assert s.graph_stats() == (0, 0)
    s.close()
    return dts, m


@pytest.mark.parametrize("precision", [rhydro.RH_PREC_FP64, rhydro.RH_PREC_MIXED])
def test_uniform_graph_replays_equal_eager_runs(lognormal5, precision):
    print(f"case uniform graph replays precision {precision}", flush=True)
    want_dts, want = _eager_uniform(lognormal5, 0.8, 4, precision)
    if precision == rhydro.RH_PREC_FP64:
        mo = clone(lognormal5)
        dto = oracle.OracleMesh(mo).run_uniform(oracle.params(1.4, 1, 2), LEVEL, DX, 0.8, 4)
        assert_same(want_dts, dto, "eager dt vs oracle")
        assert_same(want.uold, mo.uold, "eager uold vs oracle")

    m = clone(lognormal5)
    s = rhydro.Solver(gamma=1.4, slope_type=1, riemann="hllc", precision=precision)
    s.bind(m, rhydro.RH_HOST_COHERENT)
    for rep in range(3):
        m.uold[:] = lognormal5.uold
        m.unew[:] = lognormal5.unew
        dts = s.run_uniform(LEVEL, DX, 0.8, 4)
        assert s.graph_stats() == (1, rep + 1)
        assert_same(dts, want_dts, f"replay {rep} dt")
        assert_same(m.uold, want.uold, f"replay {rep} uold")
        assert_same(m.unew, want.unew, f"replay {rep} unew")
        # eager work on another context and stream between the replays
        _eager_uniform(lognormal5, 0.7, 2, precision)

    # another Courant factor: a new graph, equal to its eager run
    want_dts7, want7 = _eager_uniform(lognormal5, 0.7, 4, precision)
    m.uold[:] = lognormal5.uold
    m.unew[:] = lognormal5.unew
    dts = s.run_uniform(LEVEL, DX, 0.7, 4)
    assert s.graph_stats() == (2, 4)
    assert_same(dts, want_dts7, "cf 0.7 dt")
    assert_same(m.uold, want7.uold, "cf 0.7 uold")

    # timers on: eager launches, same result, no graph launched
    s.timers(True)
    m.uold[:] = lognormal5.uold
    m.unew[:] = lognormal5.unew
    dts = s.run_uniform(LEVEL, DX, 0.7, 4)
    assert s.graph_stats() == (2, 4)
    assert_same(m.uold, want7.uold, "timed eager uold")
    s.close()


def test_device_resident_consecutive_calls_replay(lognormal5):
    """Two calls of three steps on a device-resident mesh: the second replays
    the first call's graph on the state the first left, as eager calls do."""
    print("case device-resident replay", flush=True)
    out = {}
    for graphs in (False, True):
        m = clone(lognormal5)
        s = rhydro.Solver(gamma=1.4, slope_type=1, riemann="hllc")
        s.graphs(graphs)
        s.bind(m, rhydro.RH_DEVICE_RESIDENT)
        d1 = s.run_uniform(LEVEL, DX, 0.8, 3)
        d2 = s.run_uniform(LEVEL, DX, 0.8, 3)
        s.sync_to_host()
        out[graphs] = (np.concatenate([d1, d2]), m, s.graph_stats())
        s.close()
    assert out[False][2] == (0, 0) and out[True][2] == (1, 2)
    assert_same(out[True][0], out[False][0], "dt")
    assert_same(out[True][1].uold, out[False][1].uold, "uold")
    assert_same(out[True][1].unew, out[False][1].unew, "unew")


def _periodic_boxes():
    out = []
    for k, (nx, L) in enumerate([(1, 3), (3, 2), (1, 4), (5, 2)]):
        m = rhydro.uniform_mesh(L, nx=nx, ny=nx, nz=nx)
        rhydro.init_lognormal(m, L, 300 + k, 1.4, 0.4, 0.4)
        out.append((m, L, 1.0 / (nx * 2**L)))
    return out


def test_ensemble_graph_replays_equal_eager_and_oracle():
    print("case ensemble graph replays", flush=True)
    boxes = _periodic_boxes()
    levels = [L for _, L, _ in boxes]
    dxs = [dx for _, _, dx in boxes]

    def run(graphs, reps, other=None):
        meshes = [clone(m) for m, _, _ in boxes]
        s = rhydro.Solver(gamma=1.4, slope_type=1, riemann="hllc")
        s.graphs(graphs)
        s.bind_ensemble(meshes, rhydro.RH_HOST_COHERENT)
        res = []
        for rep in range(reps):
            for mm, (m0, _, _) in zip(meshes, boxes):
                mm.uold[:] = m0.uold
                mm.unew[:] = m0.unew
            dts = s.run_ensemble(levels, dxs, 0.8, 3)
            res.append((dts, [mm.uold.copy() for mm in meshes], [mm.unew.copy() for mm in meshes]))
            if other is not None:
                other()
        st = s.graph_stats()
        p = oracle_params(s)
        s.close()
        return res, st, p

    (want,), st, p = run(False, 1)
    assert st == (0, 0)
    for b, (m0, L, dx) in enumerate(boxes):
        mo = clone(m0)
        dto = oracle.OracleMesh(mo).run_uniform(p, L, dx, 0.8, 3)
        assert_same(want[0][:, b], dto, f"eager box {b} dt vs oracle")
        assert_same(want[1][b], mo.uold, f"eager box {b} uold vs oracle")

    got, st, _ = run(True, 3, other=lambda: run(False, 1))
    assert st == (1, 3)
    for rep, (dts, uo, un) in enumerate(got):
        assert_same(dts, want[0], f"replay {rep} dt")
        for b in range(len(boxes)):
            assert_same(uo[b], want[1][b], f"replay {rep} box {b} uold")
            assert_same(un[b], want[2][b], f"replay {rep} box {b} unew")
