"""GPU checks at the benchmark's full size: 2^21 octs (256^3 leaf cells) at
reference level 9, the lognormal state the benchmark times (seed 20261202,
gamma = 5/3, HLLC, MinMod).

The oracle cannot update the whole level in test time, so it updates the
first 16,384 octs of the list (a 64x64x32 block in the list's Morton
order).  A cell's new value is complete once its own oct and the six
neighbour octs were processed (its incoming contributions all come from
them), so the 12,600 octs of the block whose six neighbours are also in it
are compared value for value with the GPU's full-level step.  The Courant
step is a minimum over the whole level and is compared exactly too.  The
mixed form is checked against the FP64 form over every cell of the level."""
import numpy as np
import pytest

import oracle
import rhydro
from helpers import assert_same, clone

pytestmark = pytest.mark.gpu

LEVEL = 9
GAMMA = 5.0 / 3.0
DX = 1.0 / 2**LEVEL
SAMPLE_OCTS = 1 << 14


@pytest.fixture(scope="module")
def level9():
    m = rhydro.uniform_mesh(LEVEL)
    rhydro.init_lognormal(m, LEVEL, 20261202, GAMMA, 0.5, 0.5)
    return m


def _interior_cells(m, nsample):
    """Cell ids (1-based) of the octs among the first nsample of the list
    whose six same-level neighbour octs are also among them."""
    octs = np.asarray(m.level_octs(LEVEL), dtype=np.int64)
    pos = np.full(m.ngridmax + 1, -1, np.int64)
    pos[octs] = np.arange(octs.size)
    nb_cells = m.nbor.reshape(6, m.ngridmax)[:, octs[:nsample] - 1]
    nb_pos = pos[m.son[nb_cells - 1]]
    inner = ((nb_pos >= 0) & (nb_pos < nsample)).all(0)
    g = octs[:nsample][inner]
    cells = m.ncoarse + np.arange(8)[:, None] * m.ngridmax + g[None, :]
    return cells.reshape(-1), int(inner.sum())


@pytest.mark.parametrize("riemann,slope", [("hllc", 1), ("hll", 1), ("hllc", 2), ("hll", 2)])
def test_full_level_fp64_step_matches_oracle(level9, riemann, slope):
    """The benchmark workload's four variants (HLL and HLLC, slope_type 1 and
    2) at its full size."""
    mg, mo = clone(level9), clone(level9)
    s = rhydro.Solver(gamma=GAMMA, slope_type=slope, riemann=riemann)
    s.bind(mg)
    print(f"case full-size fp64 {riemann} slope {slope}: {2**21} octs at level {LEVEL}", flush=True)
    dt = s.courant(LEVEL, DX, GAMMA, 0.8)
    s.set_unew(LEVEL)
    s.godunov_fine(LEVEL, dt, DX)
    s.sync_to_host()
    s.close()

    om = oracle.OracleMesh(mo)
    assert om.courant(LEVEL, DX, GAMMA, 0.8) == dt
    om.set_unew(LEVEL)
    om.godunov_fine_sample(oracle.params(GAMMA, slope, rhydro.resolve_riemann(riemann), 1, 0, 1.0), LEVEL, dt, DX,
                           SAMPLE_OCTS)

    cells, ninner = _interior_cells(level9, SAMPLE_OCTS)
    assert ninner == 12600
    got = mg.unew.reshape(5, -1)[:, cells - 1]
    want = mo.unew.reshape(5, -1)[:, cells - 1]
    assert_same(got, want, "unew of the interior octs of the sample")
    # The step did change the state it compared.
    assert (got != level9.uold.reshape(5, -1)[:, cells - 1]).any()


def test_full_level_mixed_form_against_fp64(level9):
    """Tolerance of the mixed form after one step, per leaf cell relative to
    the FP64 form: L1 <= 1e-7 and Linf <= 1e-5 in density and total energy
    (measured about 5e-9 and 4e-7)."""
    out = {}
    dt = None
    for prec in (rhydro.RH_PREC_FP64, rhydro.RH_PREC_MIXED):
        m = clone(level9)
        s = rhydro.Solver(gamma=GAMMA, slope_type=1, riemann="hllc", precision=prec)
        s.bind(m)
        print(f"case full-size precision {prec}", flush=True)
        if dt is None:
            dt = s.courant(LEVEL, DX, GAMMA, 0.8)
        s.set_unew(LEVEL)
        s.godunov_fine(LEVEL, dt, DX)
        s.sync_to_host()
        s.close()
        cells = m.cells_of_level(LEVEL)
        out[prec] = m.unew.reshape(5, -1)[:, cells]
    ref, mix = out[rhydro.RH_PREC_FP64], out[rhydro.RH_PREC_MIXED]
    assert np.isfinite(mix).all()
    for iv, name in ((0, "density"), (4, "total energy")):
        err = np.abs(mix[iv] - ref[iv])
        l1 = err.sum() / np.abs(ref[iv]).sum()
        linf = (err / np.abs(ref[iv])).max()
        print(f"mixed vs fp64 {name}: L1 {l1:.2e} Linf {linf:.2e}", flush=True)
        assert l1 <= 1e-7, name
        assert linf <= 1e-5, name
