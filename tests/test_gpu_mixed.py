"""The mixed-precision form (FP32 face states -- the primitive conversion
of each cell on a line, slopes, trace -- and an FP32 Riemann solver and face
fluxes; FP64 flux difference, reflux, update, clamps and Courant step)
against the FP64 oracle.

The mixed form is not bit-faithful by design.  On smooth states one step
from identical input must agree with the FP64 oracle to FP32 accuracy of the
face states; the tolerances are written in each test.
"""
import numpy as np
import pytest

import oracle
import rhydro
from helpers import VARIANTS, clone, oracle_params

pytestmark = pytest.mark.gpu


def _errors(m_gpu, m_ref, level, gamma):
    cells = m_gpu.cells_of_level(level, leaf_only=True)
    a = m_gpu.state("uold")[:, cells]
    b = m_ref.state("uold")[:, cells]
    rho_a, rho_b = a[0], b[0]
    v_a, v_b = a[1:4] / rho_a, b[1:4] / rho_b
    p_a = (gamma - 1) * (a[4] - 0.5 * rho_a * (v_a ** 2).sum(0))
    p_b = (gamma - 1) * (b[4] - 0.5 * rho_b * (v_b ** 2).sum(0))
    out = {}
    for name, x, y, scale in (("rho", rho_a, rho_b, np.abs(rho_b)), ("p", p_a, p_b, np.abs(p_b)),
                              ("v", np.sqrt((v_a ** 2).sum(0)), np.sqrt((v_b ** 2).sum(0)), None)):
        d = np.abs(x - y)
        if scale is None:
            scale = np.sqrt((v_b ** 2).sum(0)).max()
        rel = d / scale
        out[name] = (float(rel.mean()), float(rel.max()))
    return out


def _run(m0, level, riemann, slope, nsteps, gamma=1.4):
    s = rhydro.Solver(gamma=gamma, slope_type=slope, riemann=riemann, precision=rhydro.RH_PREC_MIXED)
    mg, mo = clone(m0), clone(m0)
    s.bind(mg)
    dx = 1.0 / 2 ** level
    dtg = s.run_uniform(level, dx, 0.8, nsteps)
    s.sync_to_host()
    om = oracle.OracleMesh(mo)
    dto = om.run_uniform(oracle_params(s), level, dx, 0.8, nsteps)
    s.close()
    return mg, mo, dtg, dto


@pytest.mark.parametrize("riemann", ["llf", "hll", "hllc", "acoustic"])
@pytest.mark.parametrize("slope", [1, 2])
def test_mixed_single_step_smooth(riemann, slope):
    level = 5
    m = rhydro.uniform_mesh(level)
    rhydro.init_lognormal(m, level, 123, 5.0 / 3.0, 0.1, 0.1)
    print(f"case mixed {riemann} slope {slope}", flush=True)
    mg, mo, dtg, dto = _run(m, level, riemann, slope, 1, gamma=5.0 / 3.0)
    assert dtg[0] == dto[0]  # the Courant step reads only FP64 state
    err = _errors(mg, mo, level, 5.0 / 3.0)
    # FP32 face states (about 6e-8 relative) feed a flux difference scaled by
    # dt/dx ~ 0.1: per-step errors stay below 1e-6 relative.
    for k, (l1, linf) in err.items():
        assert l1 < 1e-7 and linf < 1e-5, (k, l1, linf)


@pytest.mark.parametrize("riemann", ["hll", "hllc"])
def test_mixed_ten_steps_smooth(riemann):
    """Ten steps on a smooth state, each form with its own time steps: the
    per-step FP32 flux errors accumulate to at most L1 1e-6 and Linf 1e-4
    relative in density and pressure."""
    level = 5
    m = rhydro.uniform_mesh(level)
    rhydro.init_lognormal(m, level, 321, 5.0 / 3.0, 0.1, 0.1)
    print(f"case mixed ten steps {riemann}", flush=True)
    mg, mo, dtg, dto = _run(m, level, riemann, 1, 10, gamma=5.0 / 3.0)
    assert np.isfinite(mg.uold).all()
    np.testing.assert_allclose(dtg, dto, rtol=1e-5)
    err = _errors(mg, mo, level, 5.0 / 3.0)
    print("mixed ten-step errors", err, flush=True)
    for k in ("rho", "p"):
        l1, linf = err[k]
        assert l1 < 1e-6 and linf < 1e-4, (k, l1, linf)


def test_mixed_state_at_rest_unchanged():
    level = 5
    m = rhydro.uniform_mesh(level)
    rhydro.init_regions(m, [dict(type="square", length_x=10, length_y=10, length_z=10, d=1.0, p=1.0)], 1.4)
    before = m.uold.copy()
    s = rhydro.Solver(gamma=1.4, riemann="hllc", precision=rhydro.RH_PREC_MIXED)
    s.bind(m)
    s.run_uniform(level, 1 / 32, 0.8, 3)
    s.sync_to_host()
    cells = m.cells_of_level(level)
    np.testing.assert_allclose(m.state("uold")[:, cells], before.reshape(5, -1)[:, cells], rtol=1e-6, atol=1e-7)
    s.close()


def test_mixed_sedov_errors_are_reported():
    """Sedov 32^3 over 5 steps: the run stays finite and the errors against
    the FP64 oracle are measured (the reference dynamics are chaotic, so
    multi-step errors are not small; single-step errors from identical input
    are)."""
    level = 6
    m = rhydro.uniform_mesh(level)
    rhydro.init_regions(m, rhydro.sedov_regions(level), 1.4)
    mg, mo, dtg, dto = _run(m, level, "hllc", 1, 1)
    assert np.isfinite(mg.uold).all()
    err1 = _errors(mg, mo, level, 1.4)
    print("sedov32 mixed 1-step errors", err1, flush=True)
    assert err1["rho"][0] < 1e-6


@pytest.mark.parametrize("riemann", ["llf", "hll", "hllc", "acoustic"])
def test_mixed_variants_identical(riemann):
    """Both kernel variants evaluate the same FP32/FP64 operations on the
    same inputs, so the mixed form is variant-independent to the bit."""
    level = 5
    m0 = rhydro.uniform_mesh(level)
    rhydro.init_lognormal(m0, level, 321, 1.4, 0.5, 0.5)
    outs = []
    for variant in VARIANTS:
        print(f"case mixed variant {variant} {riemann}", flush=True)
        m = clone(m0)
        s = rhydro.Solver(gamma=1.4, slope_type=2, riemann=riemann, precision=rhydro.RH_PREC_MIXED, variant=variant)
        s.bind(m)
        s.run_uniform(level, 1 / 32, 0.8, 3)
        s.sync_to_host()
        s.close()
        outs.append(m.uold.copy())
    assert (outs[0] == outs[1]).all()
