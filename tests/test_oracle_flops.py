"""Exact floating-point operation count of the reference's godunov_fine per
cell-update (oracle/flopcount.cpp: the oracle compiled with a counting
FP64 type).  The counted run must compute exactly what the plain oracle
computes; the counts it reports are the arithmetic the measurement divides
the GPU's throughput by.

Pin: six Riemann solves per cell-update with two sound speeds each give
exactly 12 square roots per cell for LLF, HLL and HLLC (a gprof profile of
the reference binary counts 6 solve_hllc calls per cell-update)."""
import numpy as np
import pytest

import oracle
import rhydro
from helpers import assert_same, clone

GAMMA = 5.0 / 3.0


def _count(level, riemann, slope_type, seed=20261202):
    m = rhydro.uniform_mesh(level)
    rhydro.init_lognormal(m, level, seed, GAMMA, 0.5, 0.5)
    dx = 1.0 / 2**level
    mo = clone(m)
    om = oracle.OracleMesh(mo)
    dt = om.courant(level, dx, GAMMA, 0.8)
    unew, counts = oracle.flop_count(m, level, GAMMA, slope_type, riemann, 1, dt, dx)
    om.set_unew(level)
    om.godunov_fine(oracle.params(GAMMA, slope_type, riemann, 1, 0, 1.0), level, dt, dx)
    cells = m.cells_of_level(level)
    assert_same(unew.reshape(5, -1)[:, cells], mo.unew.reshape(5, -1)[:, cells], "counted run against the oracle")
    return counts, cells.size


@pytest.mark.parametrize("riemann", [0, 1, 2])
@pytest.mark.parametrize("slope_type", [1, 2])
def test_counted_run_is_the_oracle(riemann, slope_type):
    counts, n = _count(5, riemann, slope_type)
    assert counts["sqrt"] == 12 * n
    assert counts["pow"] == 0
    flops = counts["add"] + counts["mul"] + counts["div"] + counts["sqrt"]
    print(f"riemann {riemann} slope {slope_type}: "
          + ", ".join(f"{k} {v / n:.1f}" for k, v in counts.items()) + f"; FP64 ops per cell {flops / n:.1f}")
    assert 2000 * n < flops < 3000 * n


def test_benchmark_state_count():
    """The benchmark's workload (HLLC, MinMod, lognormal state, gamma 5/3)
    at 32^3: 2,412.6 FP64 operations per cell-update (855.8 add/sub,
    1,329.8 mul, 215.0 div, 12 sqrt) and 326.9 comparisons.  The count is
    per cell and does not depend on the level's size beyond branch mix."""
    counts, n = _count(6, 2, 1)
    per = {k: v / n for k, v in counts.items()}
    flops = per["add"] + per["mul"] + per["div"] + per["sqrt"]
    assert abs(flops - 2412.6) < 0.5
    assert abs(per["div"] - 215.0) < 0.1
    assert per["sqrt"] == 12.0
