"""Shared helpers of the parity tests: copies of meshes, bitwise comparison
with a readable report, and the uniform-run loops of both implementations."""
import copy

import numpy as np

import oracle
import rhydro

# Every Godunov kernel variant (per-cell, oct-in-lanes, bricks); all must give
# identical results.
VARIANTS = list(range(rhydro.RH_KERNEL_VARIANTS))


def clone(mesh: rhydro.Mesh) -> rhydro.Mesh:
    m = copy.copy(mesh)
    m.arrays = {k: v.copy() for k, v in mesh.arrays.items()}
    return m


def ulp_diff(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    ia = a.view(np.int64).astype(np.int64)
    ib = b.view(np.int64).astype(np.int64)
    ia = np.where(ia < 0, np.int64(-0x8000000000000000) - ia, ia)
    ib = np.where(ib < 0, np.int64(-0x8000000000000000) - ib, ib)
    return np.abs(ia - ib)


def assert_same(got: np.ndarray, want: np.ndarray, what: str = "") -> None:
    """Equality as the reference's test harness counts it (== per value, so
    +0 and -0 agree; NaN agrees with NaN)."""
    got = np.asarray(got)
    want = np.asarray(want)
    eq = (got == want) | (np.isnan(got) & np.isnan(want))
    if not eq.all():
        bad = np.flatnonzero(~eq.reshape(-1))
        g = got.reshape(-1)[bad[:5]]
        w = want.reshape(-1)[bad[:5]]
        u = ulp_diff(got.reshape(-1)[bad], want.reshape(-1)[bad])
        raise AssertionError(f"{what}: {bad.size} of {eq.size} values differ; first at flat index {bad[:5]} "
                             f"got {g} want {w}; max ulp {u.max()}")


def oracle_params(s: rhydro.Solver) -> oracle.OrParams:
    p = s.params
    return oracle.params(p.gamma, p.slope_type, p.riemann, p.interpol_type, p.interpol_var, p.boxlen)


def step_oracle(om: oracle.OracleMesh, p: oracle.OrParams, level: int, dx: float, cf: float) -> float:
    dt = om.courant(level, dx, p.gamma, cf, p.boxlen)
    om.set_unew(level)
    om.godunov_fine(p, level, dt, dx)
    om.set_uold(level)
    return dt


def step_gpu(s: rhydro.Solver, level: int, dx: float, cf: float) -> float:
    dt = s.courant(level, dx, s.params.gamma, cf)
    s.set_unew(level)
    s.godunov_fine(level, dt, dx)
    s.set_uold(level)
    return dt
