"""The oracle's refinement criteria and flag smoothing pinned by closed forms
(TreeUpdater.cpp:328-582 flag_fine with smooth_fine :217-268, as
Simulation::initialize applies them when it refines the initial state).

A uniform level L (N = 2^(L-1) cells a side, periodic) holds a step in
density (or pressure) across the planes x = 1/2 and x = 0: rho = 1 for
x < 1/2 and 2 for x > 1/2.  flag_fine's gradient measure of a cell is
2 max(|r - c| / (r + c + 1e-10), |c - l| / (c + l + 1e-10)) along each axis,
which is 2/3 for the cells next to an interface and 0 elsewhere, so with
err_grad = 0.1 exactly the two cell layers at each of the two interfaces are
flagged: 4 N^2 cells.  Each smoothing pass (nexpand of them) flags every
cell with a flagged face neighbour, adding one layer on each side of each
band: 4 (nexpand + 1) N^2 cells, which become the octs of level L+1.  A step
of 1.05 (measure 0.049) flags nothing."""
import numpy as np
import pytest

import oracle
import rhydro

L = 5
N = 2 ** (L - 1)


def _run(kind, contrast, nexpand):
    reg = [dict(type="square", x_center=0.5, y_center=0.5, z_center=0.5, length_x=10, length_y=10, length_z=10,
                d=1.0, p=1.0),
           dict(type="square", x_center=0.75, y_center=0.5, z_center=0.5, length_x=0.5, length_y=10, length_z=10,
                d=contrast if kind == "d" else 1.0, p=contrast if kind == "p" else 1.0)]
    crit = dict(err_grad_d=0.1) if kind == "d" else dict(err_grad_p=0.1)
    return oracle.OracleAmr(levelmin=L, levelmax=L + 1, ngridmax=20000, nexpand=nexpand, regions=reg, **crit)


def _refined_x_layers(s):
    """x indices (cell units of level L) of the cells refined into level L+1."""
    nc, ng = s["ncoarse"], s["ngridmax"]
    xg = s["xg"].reshape(3, ng)
    g, xs = int(s["headl"][L + 1]), []
    while g > 0:
        xs.append(xg[0, g - 1])
        g = int(s["next"][g - 1])
    return np.floor(np.array(xs) * N).astype(int)


@pytest.mark.parametrize("kind", ["d", "p"])
@pytest.mark.parametrize("nexpand", [0, 1, 2])
def test_step_flags_the_interface_bands(kind, nexpand):
    a = _run(kind, 2.0, nexpand)
    s = a.state()
    assert s["numbl"][L] == N ** 3 // 8
    assert s["numbl"][L + 1] == 4 * (nexpand + 1) * N * N
    # Each refined x layer holds a full N x N plane, and the layers are the
    # ones within nexpand + 1 cells of the planes x = 1/2 and x = 0 (mod 1).
    xs = _refined_x_layers(s)
    layers, counts = np.unique(xs, return_counts=True)
    assert (counts == N * N).all()
    want = set()
    for k in range(nexpand + 1):
        want |= {N // 2 - 1 - k, N // 2 + k, (N - 1 - k) % N, k % N}
    assert set(layers.tolist()) == want
    a.close()


@pytest.mark.parametrize("kind", ["d", "p"])
def test_weak_step_flags_nothing(kind):
    a = _run(kind, 1.05, 1)
    assert a.state()["numbl"][L + 1] == 0
    a.close()
