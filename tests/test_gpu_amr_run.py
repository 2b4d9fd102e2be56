"""Adaptive runs with the state on the GPU (rhydro_amr.h) against the
oracle's restatement of the reference's adaptive time integration: the tree
(lists, links, geometry) and the FP64 state must be identical after the
initial refinement and after every coarse step."""
import numpy as np
import pytest

import oracle
import rhydro
from helpers import assert_same

pytestmark = pytest.mark.gpu

KEYS_INT = ("numbl", "headl", "next", "father", "son", "nbor")


def _compare(g, o, what):
    for k in KEYS_INT:
        if not np.array_equal(g[k], o[k]):
            bad = np.flatnonzero(g[k] != o[k])[:5]
            raise AssertionError(f"{what}: {k} differs at {bad}: gpu {g[k][bad]} oracle {o[k][bad]}")
    assert_same(g["xg"], o["xg"], f"{what}: xg")
    assert_same(g["uold"], o["uold"], f"{what}: uold")
    assert_same(g["unew"], o["unew"], f"{what}: unew")
    assert g["t"] == o["t"] and g["nstep"] == o["nstep"], what
    assert g["cell_updates"] == o["cell_updates"], what


def _pair(**kw):
    return rhydro.AmrRun(**kw), oracle.OracleAmr(**{k: (rhydro.resolve_riemann(v) if k == "riemann" else v)
                                                   for k, v in kw.items()})


@pytest.mark.parametrize("riemann", ["hllc", "llf"])
def test_sedov_amr_run(riemann):
    kw = dict(levelmin=4, levelmax=6, ngridmax=6000, err_grad_d=0.1, err_grad_p=0.1, riemann=riemann,
              regions=rhydro.sedov_regions(4))
    print(f"case amr run sedov {riemann}", flush=True)
    g, o = _pair(**kw)
    _compare(g.state(), o.state(), "initial mesh")
    assert g.state()["numbl"][5] > 0
    for k in range(4):
        g.step(1)
        assert o.step(1) == 0
        _compare(g.state(), o.state(), f"coarse step {k + 1}")
    g.close()
    o.close()


def test_sedov_amr_run_four_levels():
    """Four refined levels and six coarse steps: the blast grows, so level
    lists gain and lose octs at every depth while the level tables of the
    lists not next to a change are kept from one step to the next."""
    kw = dict(levelmin=4, levelmax=7, ngridmax=20000, err_grad_d=0.1, err_grad_p=0.1, riemann="hllc",
              regions=rhydro.sedov_regions(4))
    print("case amr run sedov levels 4-7", flush=True)
    g, o = _pair(**kw)
    _compare(g.state(), o.state(), "initial mesh")
    for k in range(6):
        g.step(1)
        assert o.step(1) == 0
        _compare(g.state(), o.state(), f"coarse step {k + 1}")
    assert g.state()["numbl"][7] > 0
    g.close()
    o.close()


def test_kelvin_helmholtz_amr_run():
    kw = dict(levelmin=5, levelmax=6, ngridmax=4000, err_grad_d=0.2, slope_type=1, riemann="hllc", ic="kh",
              kh_amp=0.05)
    print("case amr run kh", flush=True)
    g, o = _pair(**kw)
    _compare(g.state(), o.state(), "initial mesh")
    g.step(3)
    assert o.step(3) == 0
    _compare(g.state(), o.state(), "after 3 coarse steps")
    g.close()
    o.close()


def test_uniform_run_through_the_adaptive_driver():
    kw = dict(levelmin=5, levelmax=5, ngridmax=int(rhydro.lib().rh_uniform_ngrid(1, 1, 1, 5)),
              regions=rhydro.sedov_regions(5))
    print("case amr run uniform", flush=True)
    g, o = _pair(**kw)
    g.step(3)
    o.step(3)
    _compare(g.state(), o.state(), "uniform")
    g.close()


def test_mixed_form_amr_run_stays_finite():
    print("case amr run mixed", flush=True)
    g = rhydro.AmrRun(levelmin=4, levelmax=6, ngridmax=6000, err_grad_d=0.1, err_grad_p=0.1,
                      regions=rhydro.sedov_regions(4), precision=rhydro.RH_PREC_MIXED)
    g.step(3)
    s = g.state()
    assert np.isfinite(s["uold"]).all() and s["numbl"][6] > 0
    g.close()


def test_exhausted_grids_are_reported():
    """289 grids hold the initial mesh; refining further needs more than 300."""
    print("case amr run exhausted", flush=True)
    g = rhydro.AmrRun(levelmin=4, levelmax=6, ngridmax=300, err_grad_d=0.1, err_grad_p=0.1,
                      regions=rhydro.sedov_regions(4))
    with pytest.raises(rhydro.RhError):
        g.step(4)
    g.close()


def test_sedov_amr_run_at_benchmark_levels():
    """The benchmark's adaptive Sedov run at its own levels (64^3 base at
    reference level 7, refined to level 10, err_grad 0.1): the initial mesh
    and the first two coarse steps (the blast reaches level 10 in the first)
    identical to the oracle's restatement."""
    kw = dict(levelmin=7, levelmax=10, ngridmax=300000, err_grad_d=0.1, err_grad_p=0.1, riemann="hllc",
              regions=rhydro.sedov_regions(7))
    print("case amr run sedov levels 7-10", flush=True)
    g, o = _pair(**kw)
    _compare(g.state(), o.state(), "initial mesh")
    for k in range(2):
        g.step(1)
        assert o.step(1) == 0
        _compare(g.state(), o.state(), f"coarse step {k + 1}")
    assert g.state()["numbl"][10] > 0
    g.close()
    o.close()


def test_kelvin_helmholtz_amr_run_scaled():
    """The adaptive Kelvin-Helmholtz slab of the benchmark two levels
    coarser (32^3 base at reference level 6, refined to level 9 on density
    gradients, err_grad_d 0.2): one coarse step (8 fine steps on the finest
    level) identical to the oracle's restatement, in the FP64 form."""
    kw = dict(levelmin=6, levelmax=9, ngridmax=250000, err_grad_d=0.2, ic="kh", kh_amp=0.01, riemann="hllc")
    print("case amr run kh levels 6-9", flush=True)
    g, o = _pair(**kw)
    _compare(g.state(), o.state(), "initial mesh")
    g.step(1)
    assert o.step(1) == 0
    _compare(g.state(), o.state(), "coarse step 1")
    assert g.state()["numbl"][9] > 0
    g.close()
    o.close()


def test_kelvin_helmholtz_initial_mesh_at_benchmark_levels():
    """The benchmark's adaptive Kelvin-Helmholtz slab at its own levels
    (128^3 base at reference level 8, refined to level 11 on density
    gradients, err_grad_d 0.2; 2,097,152 octs on level 11): the initial
    refinement -- every level list in order, the links, the geometry and the
    FP64 state, restricted upward -- identical to the oracle's restatement.
    (A coarse step at this size takes the oracle minutes; the same step is
    compared two levels coarser above.)"""
    kw = dict(levelmin=8, levelmax=11, ngridmax=4500000, err_grad_d=0.2, ic="kh", kh_amp=0.01, riemann="hllc")
    print("case amr run kh levels 8-11, initial mesh", flush=True)
    g = rhydro.AmrRun(**kw)
    gs = g.state()
    g.close()
    print("gpu run initialised", flush=True)
    o = oracle.OracleAmr(**{k: (rhydro.resolve_riemann(v) if k == "riemann" else v) for k, v in kw.items()})
    print("oracle initialised", flush=True)
    _compare(gs, o.state(), "initial mesh")
    assert gs["numbl"][11] == 2097152
    o.close()
