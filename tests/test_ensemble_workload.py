"""The shock-tube ensemble workload (host side): box sizes, cell widths and
initial states."""
import numpy as np

import rhydro
from amr_meshes import cell_centres

SIDES = {16, 20, 24, 28, 32, 40, 48, 56, 64}


def test_quantised_sides_are_the_reference_realisable_ones():
    table = rhydro._quantised_sides(16, 64)
    assert {t[0] for t in table} == SIDES
    for side, nx, L in table:
        assert side == nx * 2 ** (L - 1) and nx in (1, 3, 5, 7) and L >= 2


def test_ensemble_is_reproducible_and_spans_the_range():
    a = rhydro.shock_tube_ensemble(24, seed=5)
    b = rhydro.shock_tube_ensemble(24, seed=5)
    c = rhydro.shock_tube_ensemble(24, seed=6)
    sides = [bx.nx * 2 ** (bx.level - 1) for bx in a]
    assert set(sides) <= SIDES and len(set(sides)) >= 5
    for x, y in zip(a, b):
        assert (x.nx, x.level) == (y.nx, y.level)
        np.testing.assert_array_equal(x.mesh.uold, y.mesh.uold)
    assert any(not np.array_equal(x.normal, y.normal) for x, y in zip(a, c))


def test_box_geometry_and_states():
    for bx in rhydro.shock_tube_ensemble(6, seed=11):
        m = bx.mesh
        side = bx.nx * 2 ** (bx.level - 1)
        assert m.nx == m.ny == m.nz == bx.nx
        assert len(m.level_octs(bx.level)) * 8 == side ** 3
        assert bx.dx == 1.0 / (bx.nx * 2 ** bx.level)  # the reference's dx = boxlen / (nx * 2^L)
        assert abs(np.linalg.norm(bx.normal) - 1.0) < 1e-12
        assert 0.1 <= bx.left[0] <= 1 and 0.1 <= bx.left[4] <= 1
        cells, cen = cell_centres(m, bx.level)
        s = (cen - 0.5) @ bx.normal
        u = m.state("uold")[:, cells - 1]
        left = s < 0
        np.testing.assert_array_equal(u[0, left], bx.left[0])
        np.testing.assert_array_equal(u[0, ~left], bx.right[0])
        np.testing.assert_array_equal(u[4, left], bx.left[4] / (1.4 - 1.0))
        np.testing.assert_array_equal(u[1:4], 0.0)
        np.testing.assert_array_equal(m.unew, m.uold)


def test_axis_aligned_tube_splits_at_the_plane():
    m = rhydro.uniform_mesh(4, nx=3, ny=3, nz=3)
    rhydro.init_shock_tube(m, [1, 0, 0], [0.5, 0.5, 0.5], [1, 0, 0, 0, 1], [0.125, 0, 0, 0, 0.1], 1.4)
    cells, cen = cell_centres(m, 4)
    rho = m.state("uold")[0, cells - 1]
    np.testing.assert_array_equal(rho, np.where(cen[:, 0] < 0.5, 1.0, 0.125))
    assert (cen[:, 0] < 0.5).sum() == (cen[:, 0] > 0.5).sum()


def test_mixed_ensemble_sedov_boxes():
    """box_ensemble: about half the boxes are Sedov blasts with the energy
    deposited in the 8 cells around the centre (p = (gamma-1) E / (8 h^3),
    Initializer.cpp:87-94), the rest rho = 1, p = 1e-5 at rest; the
    shock-tube boxes are drawn as before and sedov_share = 0 gives exactly
    shock_tube_ensemble's boxes."""
    boxes = rhydro.box_ensemble(32, seed=21)
    kinds = [bx.kind for bx in boxes]
    assert 8 <= kinds.count("sedov") <= 24 and "sod" in kinds
    gamma = 1.4
    for bx in boxes:
        if bx.kind != "sedov":
            continue
        assert 0.1 <= bx.energy <= 1.0
        m = bx.mesh
        h = 1.0 / (bx.nx * 2 ** (bx.level - 1))
        cells, cen = cell_centres(m, bx.level)
        u = m.state("uold")[:, cells - 1]
        hot = ((cen - 0.5) ** 2).sum(1) < (1.5 * h) ** 2
        assert hot.sum() == 8
        np.testing.assert_array_equal(u[0], 1.0)
        np.testing.assert_array_equal(u[1:4], 0.0)
        p_hot = (gamma - 1.0) * bx.energy / (8 * h ** 3)
        np.testing.assert_allclose(u[4, hot] * (gamma - 1.0), p_hot, rtol=1e-14)
        np.testing.assert_allclose(u[4, ~hot] * (gamma - 1.0), 1e-5, rtol=1e-14)
        # the deposited energy is E (up to the background)
        np.testing.assert_allclose(u[4, hot].sum() * h ** 3, bx.energy, rtol=1e-12)
    a = rhydro.box_ensemble(12, seed=5, sedov_share=0.0)
    b = rhydro.shock_tube_ensemble(12, seed=5)
    for x, y in zip(a, b):
        assert x.kind == y.kind == "sod"
        np.testing.assert_array_equal(x.mesh.uold, y.mesh.uold)
