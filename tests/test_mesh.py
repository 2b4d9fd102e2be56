"""The driver's uniform mesh against the reference's construction sequence
(restated in the oracle), and geometric invariants of the octree."""
import numpy as np
import pytest

import oracle
import rhydro


@pytest.mark.parametrize("nx,ny,nz,level", [(1, 1, 1, 2), (1, 1, 1, 3), (1, 1, 1, 5), (2, 1, 1, 3), (2, 3, 2, 4),
                                            (3, 3, 3, 3)])
def test_builder_matches_reference_sequence(nx, ny, nz, level):
    m = rhydro.uniform_mesh(level, nx, ny, nz)
    a, used = oracle.build_uniform(nx, ny, nz, level, level, m.ngridmax)
    assert used == m.ngridmax
    for k in ("father", "son", "nbor", "next", "prev", "headl", "taill", "numbl"):
        assert np.array_equal(a[k], m.arrays[k]), k


def test_builder_with_spare_grids_keeps_free_list():
    m = rhydro.uniform_mesh(3, ngridmax=20, nlevelmax=5)
    a, used = oracle.build_uniform(1, 1, 1, 3, 5, 20)
    assert used == 9
    for k in ("father", "son", "nbor", "headl", "taill", "numbl"):
        assert np.array_equal(a[k], m.arrays[k]), k
    # level lists, then the free list from grid 10 on
    assert m.next[8] == 0 and m.next[9] == 11 and m.next[19] == 0


def _centres(m, level):
    octs = np.array(m.level_octs(level))
    h = 1.0 / (m.nx * 2 ** (level - 1))
    xg = m.xg.reshape(3, -1)[:, octs - 1] / m.nx
    out = {}
    for ic0 in range(8):
        off = np.array([(ic0 & 1) - 0.5, ((ic0 >> 1) & 1) - 0.5, ((ic0 >> 2) & 1) - 0.5])[:, None] * h
        for k, g in enumerate(octs):
            out[m.ncoarse + ic0 * m.ngridmax + g] = (xg[:, k] + off[:, 0])
    return out, h


@pytest.mark.parametrize("level", [3, 4])
def test_face_neighbours_are_adjacent_cells(level):
    """Every face neighbour the reference's lookup rules give (son of nbor
    cell, sibling across the oct) is the geometrically adjacent cell."""
    m = rhydro.uniform_mesh(level)
    cen, h = _centres(m, level)
    jjj = lambda ic0, d: ic0 ^ (1 << d)
    for g in m.level_octs(level):
        for ic0 in range(8):
            c = m.ncoarse + ic0 * m.ngridmax + g
            for d in range(3):
                for s in range(2):
                    cross = ((ic0 >> d) & 1) == s
                    if cross:
                        nb = m.nbor[(2 * d + s) * m.ngridmax + g - 1]
                        n = m.ncoarse + jjj(ic0, d) * m.ngridmax + m.son[nb - 1]
                    else:
                        n = m.ncoarse + jjj(ic0, d) * m.ngridmax + g
                    delta = (cen[n] - cen[c]) % 1.0
                    want = np.zeros(3)
                    want[d] = h if s == 1 else 1.0 - h
                    assert np.allclose(delta, want), (g, ic0, d, s)


def test_level_list_order_is_morton():
    """Processing order (list position x 8 + cell) of a uniform level is the
    Morton order of the cell coordinates."""
    level = 4
    m = rhydro.uniform_mesh(level)
    cen, h = _centres(m, level)
    octs = m.level_octs(level)
    keys = []
    for pos, g in enumerate(octs):
        for ic0 in range(8):
            c = m.ncoarse + ic0 * m.ngridmax + g
            ijk = np.floor(cen[c] / h).astype(int)
            code = 0
            for bit in range(level - 1):
                for d in range(3):
                    code |= ((ijk[d] >> bit) & 1) << (3 * bit + d)
            keys.append((pos * 8 + ic0, code))
    keys.sort()
    assert [k[1] for k in keys] == sorted(k[1] for k in keys)
    assert len(keys) == 8 * len(octs) == 8 ** (level - 1)
