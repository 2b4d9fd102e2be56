"""The oracle against the committed vectors of the reference's own Riemann
solvers, slope limiter and sound speed (golden/ref_kernel_vectors.npz,
written by golden/make_ref_vectors.py from the reference's sources compiled
in place), bit for bit.  Unlike test_oracle_reference.py this needs no
reference tree: the vectors travel with the repository."""
import os

import numpy as np
import pytest

import oracle

V = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_kernel_vectors.npz"))
KINDS = {0: "llf", 1: "hll", 2: "hllc", 3: "exact", 4: "acoustic"}


@pytest.mark.parametrize("kind", sorted(KINDS))
@pytest.mark.parametrize("tag,gamma", [("g14", 1.4), ("g53", 5.0 / 3.0)])
def test_riemann_vectors(kind, tag, gamma):
    ql, qr, f = (V[f"riemann{kind}_{tag}_{k}"] for k in ("ql", "qr", "f"))
    for i in range(len(ql)):
        got = oracle.riemann(kind, ql[i], qr[i], gamma)
        assert got.tobytes() == f[i].tobytes(), (KINDS[kind], gamma, i, got, f[i])


@pytest.mark.parametrize("slope_type", range(9))
def test_slope_vectors(slope_type):
    L = oracle.lib()
    trip, nus, ref = V["slope_in"], V["slope_nu"], V[f"slope{slope_type}"]
    for i, (a, b, c) in enumerate(trip):
        for j, nu in enumerate(nus):
            got = np.float64(L.or_compute_slope(a, b, c, slope_type, 1.5, nu))
            assert got.tobytes() == np.float64(ref[i, j]).tobytes(), (slope_type, a, b, c, nu)


def test_cs2_vectors():
    L = oracle.lib()
    for k, g in enumerate(V["cs2_gamma"]):
        for i in range(V["cs2_d"].size):
            got = np.float64(L.or_get_cs2(V["cs2_d"][i], V["cs2_p"][i], g))
            assert got.tobytes() == np.float64(V["cs2"][k, i]).tobytes()


def test_vectors_cover_edge_cases():
    """Floors, equal and mirror states and supersonic flow are in the set."""
    ql, qr = V["riemann2_g14_ql"], V["riemann2_g14_qr"]
    assert (ql[:, 0] < 0).any() and (ql[:, 4] < 0).any() and (np.abs(ql[:, 1]) > 1e3).any()
    assert (ql[16:22] == qr[16:22]).all()
