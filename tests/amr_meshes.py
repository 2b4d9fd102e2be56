"""AMR test meshes: a uniform level refined inside spheres, the way the
reference's make_grid_fine creates octs (restated in the oracle)."""
import numpy as np

import oracle
import rhydro


def cell_centres(m, level):
    octs = np.array(m.level_octs(level))
    h = 1.0 / (m.nx * 2 ** (level - 1))
    xg = m.xg.reshape(3, -1)[:, octs - 1] / m.nx
    cells, cen = [], []
    for g_i, g in enumerate(octs):
        for ic0 in range(8):
            off = np.array([(ic0 & 1) - 0.5, ((ic0 >> 1) & 1) - 0.5, ((ic0 >> 2) & 1) - 0.5]) * h
            cells.append(m.ncoarse + ic0 * m.ngridmax + g)
            cen.append(xg[:, g_i] + off)
    return np.array(cells), np.array(cen)


def refined_mesh(base_level=4, spheres=((0.4, 0.5, 0.6, 0.3),), extra_grids=4000, seed=7, gamma=1.4,
                 riemann=2, interpol_type=1, interpol_var=0, extra_levels=0):
    """Uniform `base_level`, then each sphere (x, y, z, r) refines the leaf
    cells of the next level whose centres lie inside it.  `extra_levels`
    leaves room in the level lists for later refinements."""
    ng = int(rhydro.lib().rh_uniform_ngrid(1, 1, 1, base_level))
    nlev = base_level + len(spheres) + extra_levels
    m = rhydro.uniform_mesh(base_level, ngridmax=ng + extra_grids, nlevelmax=nlev)
    rhydro.init_lognormal(m, base_level, seed, gamma, 0.3, 0.3)
    p = oracle.params(gamma, 1, riemann, interpol_type, interpol_var)
    for k, (x, y, z, r) in enumerate(spheres):
        lev = base_level + k
        cells, cen = cell_centres(m, lev)
        inside = ((cen - np.array([x, y, z])) ** 2).sum(1) < r * r
        leaf = m.son[cells - 1] == 0
        oracle.refine(m, p, lev, cells[inside & leaf])
    m.unew[:] = m.uold
    return m
