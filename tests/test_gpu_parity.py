"""GPU parity of the FP64 path against the oracle, through the C-ABI.

Every comparison is value-exact (== per double): the FP64 form must
reproduce the reference's arithmetic operation for operation.
"""
import numpy as np
import pytest

import oracle
import rhydro
from helpers import VARIANTS, assert_same, clone, oracle_params, step_gpu, step_oracle

pytestmark = pytest.mark.gpu

RIEMANNS = ["llf", "hll", "hllc", "acoustic"]


def _random_mesh(level, seed, gamma=5.0 / 3.0, sigma=0.5, amp=0.5):
    m = rhydro.uniform_mesh(level)
    rhydro.init_lognormal(m, level, seed, gamma, sigma, amp)
    return m


def _compare_steps(m0, solver_kw, level, nsteps, cf=0.8, mode=rhydro.RH_DEVICE_RESIDENT):
    print(f"case level={level} nsteps={nsteps} {solver_kw} mode={mode}", flush=True)
    s = rhydro.Solver(**solver_kw)
    p = oracle_params(s)
    mg, mo = clone(m0), clone(m0)
    s.bind(mg, mode)
    om = oracle.OracleMesh(mo)
    dx = 1.0 / 2 ** level
    for k in range(nsteps):
        dto = step_oracle(om, p, level, dx, cf)
        dtg = step_gpu(s, level, dx, cf)
        assert dtg == dto, f"step {k + 1}: dt gpu {dtg!r} oracle {dto!r}"
    s.sync_to_host()
    assert_same(mg.uold, mo.uold, "uold")
    assert_same(mg.unew, mo.unew, "unew")
    s.close()


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("riemann", RIEMANNS)
@pytest.mark.parametrize("slope_type", range(9))
def test_step_riemann_slope(riemann, slope_type, variant):
    m = _random_mesh(4, seed=1000 + slope_type)
    _compare_steps(m, dict(gamma=5.0 / 3.0, slope_type=slope_type, riemann=riemann, variant=variant), 4, 2)


@pytest.mark.parametrize("interpol_type", range(5))
def test_step_interpol_type(interpol_type):
    m = _random_mesh(4, seed=77)
    _compare_steps(m, dict(gamma=1.4, slope_type=2, riemann="hllc", interpol_type=interpol_type), 4, 2)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("level", [2, 3, 5])
def test_small_and_degenerate_levels(level, variant):
    # level 2 is a single oct whose every neighbour is itself (periodic).
    m = _random_mesh(level, seed=5 + level)
    _compare_steps(m, dict(gamma=1.4, slope_type=1, riemann="hllc", variant=variant), level, 3)


def test_multi_coarse_cells():
    m = rhydro.uniform_mesh(4, nx=2, ny=3, nz=1)
    rhydro.init_lognormal(m, 4, 99, 1.4)
    _compare_steps(m, dict(gamma=1.4, slope_type=2, riemann="hll"), 4, 2)


def test_host_coherent_mode():
    m = _random_mesh(4, seed=3)
    _compare_steps(m, dict(gamma=1.4, slope_type=1, riemann="hllc"), 4, 2, mode=rhydro.RH_HOST_COHERENT)


def test_sedov_32_trajectory():
    """Sedov, literal level 6 (32^3), 10 steps, HLLC/MinMod."""
    level = 6
    m = rhydro.uniform_mesh(level)
    rhydro.init_regions(m, rhydro.sedov_regions(level), 1.4)
    _compare_steps(m, dict(gamma=1.4, slope_type=1, riemann="hllc"), level, 10)


@pytest.mark.parametrize("variant", VARIANTS)
def test_sedov_64_run_uniform_20_steps(variant):
    """Config 1 at its stated size (64^3, reference level 7), 20 steps
    through rh_run_uniform, against the oracle step by step."""
    level = 7
    m = rhydro.uniform_mesh(level)
    rhydro.init_regions(m, rhydro.sedov_regions(level), 1.4)
    mg, mo = clone(m), clone(m)
    print(f"case sedov64 run_uniform 20 steps variant {variant}", flush=True)
    s = rhydro.Solver(gamma=1.4, slope_type=1, riemann="hllc", variant=variant)
    s.bind(mg)
    dx = 1.0 / 2 ** level
    dts_g = s.run_uniform(level, dx, 0.8, 20)
    s.sync_to_host()
    om = oracle.OracleMesh(mo)
    dts_o = om.run_uniform(oracle_params(s), level, dx, 0.8, 20)
    assert_same(dts_g, dts_o, "dt sequence")
    assert_same(mg.uold, mo.uold, "uold")
    assert_same(mg.unew, mo.unew, "unew")
    s.close()


def test_courant_random_states():
    m = _random_mesh(5, seed=11, sigma=2.0, amp=3.0)
    s = rhydro.Solver(gamma=5.0 / 3.0)
    s.bind(m)
    om = oracle.OracleMesh(clone(m))
    for dx, cf in [(1 / 64, 0.8), (1e-3, 0.5), (0.37, 1e-3)]:
        print(f"case courant dx={dx} cf={cf}", flush=True)
        assert s.courant(5, dx, 5.0 / 3.0, cf) == om.courant(5, dx, 5.0 / 3.0, cf)
    s.close()


def test_courant_with_gravity():
    m = _random_mesh(4, seed=12)
    f = np.random.default_rng(0).normal(0, 3.0, 3 * m.ncell)
    s = rhydro.Solver(gamma=1.4)
    s.bind(m, f=f)
    om = oracle.OracleMesh(clone(m), f=f)
    print("case courant gravity", flush=True)
    assert s.courant(4, 1 / 32, 1.4, 0.8) == om.courant(4, 1 / 32, 1.4, 0.8)
    s.close()


@pytest.mark.parametrize("interpol_var", [0, 1, 2])
@pytest.mark.parametrize("interpol_type", [0, 1, 2, 3, 4])
def test_interpol_hydro(interpol_var, interpol_type):
    rng = np.random.default_rng(interpol_var * 10 + interpol_type)
    n = 64
    u1 = np.zeros((n, 7, 64))
    rho = rng.lognormal(0, 0.7, (n, 7))
    v = rng.normal(0, 1, (n, 7, 3))
    p = rng.lognormal(0, 0.7, (n, 7))
    u1[:, :, 0] = rho
    u1[:, :, 1:4] = rho[..., None] * v
    u1[:, :, 4] = p / 0.4 + 0.5 * rho * (v ** 2).sum(-1)
    print(f"case interpol var={interpol_var} type={interpol_type}", flush=True)
    s = rhydro.Solver(gamma=1.4, interpol_type=interpol_type, interpol_var=interpol_var)
    got = s.interpol(u1)
    po = oracle.params(gamma=1.4, interpol_type=interpol_type, interpol_var=interpol_var)
    for o in range(n):
        want = oracle.interpol(po, u1[o])
        assert_same(got[o, :, :5], want[:, :5], f"oct {o}")
    s.close()


def test_state_at_rest_is_unchanged_large():
    """Uniform gas at rest on a 128^3 level stays exactly unchanged."""
    level = 8
    m = rhydro.uniform_mesh(level)
    rhydro.init_regions(m, [dict(type="square", length_x=10, length_y=10, length_z=10, d=1.3, p=0.7)], 1.4)
    before = m.uold.copy()
    s = rhydro.Solver(gamma=1.4, slope_type=2, riemann="hllc")
    s.bind(m)
    print("case at-rest 128^3", flush=True)
    s.run_uniform(level, 1 / 2 ** level, 0.8, 3)
    s.sync_to_host()
    cells = m.cells_of_level(level)
    for iv in range(5):
        a = m.state("uold")[iv, cells]
        b = before.reshape(5, -1)[iv, cells]
        assert_same(a, b, f"var {iv}")
    s.close()


def test_errors_are_reported():
    with pytest.raises(rhydro.RhError):
        rhydro.Solver(riemann="exact")
    s = rhydro.Solver()
    with pytest.raises(rhydro.RhError):
        s.godunov_fine(3, 1e-3, 1e-2)  # nothing bound
    m = _random_mesh(3, seed=1)
    s.bind(m)
    with pytest.raises(rhydro.RhError):
        s.godunov_fine(0, 1e-3, 1e-2)  # coarse level
    with pytest.raises(rhydro.RhError):
        s.godunov_fine(7, 1e-3, 1e-2)  # beyond nlevelmax
    s.close()


def test_autotune_keeps_state_and_results():
    level = 5
    m = _random_mesh(level, seed=42)
    mo = clone(m)
    s = rhydro.Solver(gamma=5.0 / 3.0, slope_type=1, riemann="hllc")
    s.bind(m)
    print("case autotune", flush=True)
    best, ms = s.autotune(level, 1e-3, 1 / 32, reps=2)
    assert best in range(rhydro.RH_KERNEL_VARIANTS) and s.variant() == best
    assert ms.shape == (rhydro.RH_KERNEL_VARIANTS,) and (ms > 0).all()
    s.sync_to_host()
    assert_same(m.unew, mo.unew, "unew restored")
    om = oracle.OracleMesh(mo)
    p = oracle_params(s)
    for _ in range(2):
        assert step_gpu(s, level, 1 / 32, 0.8) == step_oracle(om, p, level, 1 / 32, 0.8)
    s.sync_to_host()
    assert_same(m.uold, mo.uold, "uold after tuned steps")
    s.close()


def test_empty_levels_are_no_ops():
    """Level lists without octs (above the built levels): the Courant step is
    the initial value, set_unew/godunov_fine/set_uold change nothing, as in
    the oracle."""
    m = rhydro.uniform_mesh(3, nlevelmax=5)
    rhydro.init_lognormal(m, 3, 3, 1.4)
    mo = clone(m)
    before = m.uold.copy()
    s = rhydro.Solver(gamma=1.4, slope_type=1, riemann="hllc")
    s.bind(m)
    om = oracle.OracleMesh(mo)
    p = oracle_params(s)
    print("case empty levels 4 and 5", flush=True)
    for level in (4, 5):
        dtg = step_gpu(s, level, 1.0 / 2 ** level, 0.8)
        dto = step_oracle(om, p, level, 1.0 / 2 ** level, 0.8)
        assert dtg == dto, (level, dtg, dto)
    s.sync_to_host()
    assert_same(m.uold, before, "uold unchanged")
    assert_same(m.uold, mo.uold, "uold")
    assert_same(m.unew, mo.unew, "unew")
    s.close()


@pytest.mark.parametrize("riemann", ["llf", "hllc"])
def test_nonfinite_and_degenerate_cells(riemann):
    """NaN and infinite densities and energies, zero and negative densities,
    negative pressures: the GPU path propagates them exactly as the oracle
    (NaN where the oracle has NaN, the same floors and clamps elsewhere), and
    the Courant step skips NaN candidates as std::min does."""
    level = 4
    m = _random_mesh(level, seed=99, gamma=1.4)
    cells = m.cells_of_level(level)
    u = m.state("uold")
    rng = np.random.default_rng(7)
    pick = rng.choice(cells.size, 24, replace=False)
    c = cells[pick]
    u[0, c[0:3]] = np.nan
    u[4, c[3:6]] = np.inf
    u[0, c[6:9]] = 0.0
    u[0, c[9:12]] = -1.0
    u[4, c[12:15]] = -5.0
    u[1, c[15:18]] = -np.inf
    u[0, c[18:21]] = 1e-300
    u[4, c[21:24]] = np.nan
    m.unew[:] = m.uold
    _compare_steps(m, dict(gamma=1.4, slope_type=1, riemann=riemann), level, 2)


def test_amr_run_reports_exhausted_grids():
    """A run whose refinement needs more grids than ngridmax stops with an
    error (the driver does not resize), not a fault."""
    print("case amr run ngridmax too small", flush=True)
    with pytest.raises(rhydro.RhError):
        rhydro.AmrRun(levelmin=4, levelmax=6, ngridmax=80, err_grad_d=0.1, err_grad_p=0.1,
                      regions=rhydro.sedov_regions(4))


@pytest.mark.parametrize("precision", [rhydro.RH_PREC_FP64, rhydro.RH_PREC_MIXED])
def test_run_uniform_exchanged_buffers_match_copies(precision):
    """On a uniform periodic level rh_run_uniform with the oct-in-lanes
    kernels exchanges uold and unew between steps and takes the next Courant
    step from the Godunov launch; the per-cell kernel runs the copying
    sequence (set_unew, godunov_fine, fused set_uold/Courant/set_unew).  Both
    leave the same uold, unew and time steps after 1, 2 and 3 steps."""
    level = 4
    m0 = _random_mesh(level, seed=77)
    for n in (1, 2, 3):
        out = []
        for variant in VARIANTS:
            m = clone(m0)
            s = rhydro.Solver(gamma=5.0 / 3.0, slope_type=1, riemann="hllc", precision=precision, variant=variant)
            s.bind(m)
            print(f"case exchanged buffers precision {precision} steps {n} variant {variant}", flush=True)
            dts = s.run_uniform(level, 1 / 16, 0.8, n)
            s.sync_to_host()
            s.close()
            out.append((np.array(dts), m.uold.copy(), m.unew.copy()))
        for v in VARIANTS[1:]:
            assert_same(out[v][0], out[0][0], f"dt, {n} steps, variant {v}")
            assert_same(out[v][1], out[0][1], f"uold, {n} steps, variant {v}")
            assert_same(out[v][2], out[0][2], f"unew, {n} steps, variant {v}")


def test_set_unew_oct_lanes_equals_set_unew():
    """The counter-calibration copy (the Godunov kernel's lane mapping) leaves
    unew exactly as set_unew does, on every level of an adaptive mesh (lists
    of 1 to some thousands of octs, not a multiple of eight) and on a uniform
    level."""
    from amr_meshes import refined_mesh

    meshes = [(refined_mesh(4, spheres=((0.4, 0.5, 0.6, 0.3), (0.45, 0.5, 0.55, 0.15))), range(1, 7)),
              (_random_mesh(5, 11), [5])]
    for m0, levels in meshes:
        for lev in levels:
            print(f"case set_unew_oct_lanes level {lev}", flush=True)
            out = []
            for lanes in (False, True):
                m = clone(m0)
                m.unew[:] = -7.0
                s = rhydro.Solver(gamma=1.4)
                s.bind(m, rhydro.RH_HOST_COHERENT)
                if lanes:
                    s.set_unew_oct_lanes(lev)
                else:
                    s.set_unew(lev)
                s.close()
                out.append(m.unew.copy())
            assert_same(out[1], out[0], f"level {lev} unew")
            assert (out[1] != -7.0).any() or len(m0.level_octs(lev)) == 0
