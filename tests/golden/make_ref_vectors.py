"""Writes tests/golden/ref_kernel_vectors.npz: input/output vectors of the
reference's own Riemann solvers (LLF, HLL, HLLC, exact, acoustic), slope
limiter (types 0-8) and sound speed, computed by the reference's sources
compiled in place (oracle/_ref/libref_kernels.so: RiemannSolver.cpp,
SlopeLimiter.cpp, EquationOfState.cpp).  Run from the repository root where
the reference tree is present (`make -C oracle` builds oracle/_ref first);
the fixture then pins the oracle on hosts without the reference.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

N = 48  # states per solver and gamma


def states(rng, n):
    """Primitive states with the awkward cases mixed in: negative density or
    pressure (floors), zero and strongly supersonic velocity, tiny values."""
    q = np.empty((n, 5))
    q[:, 0] = rng.lognormal(0, 2, n)
    q[:, 1:4] = rng.normal(0, 3, (n, 3))
    q[:, 4] = rng.lognormal(0, 3, n)
    q[0:3, 0] *= -1
    q[3:6, 4] *= -1
    q[6:9, 1] = 0.0
    q[9:12, 1] *= 1e4
    q[12:14, 0] = 1e-12
    q[14:16, 4] = 1e-30
    return q


def main():
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref is not built: run `make -C oracle` with the reference tree present")
    rng = np.random.default_rng(20261202)
    out = {}
    for kind in (0, 1, 2, 3, 4):
        ql, qr = states(rng, N), states(rng, N)
        qr[16:22] = ql[16:22]  # equal states
        qr[22:28] = ql[22:28] * np.array([1, -1, 1, 1, 1])  # mirror states
        for tag, gamma in (("g14", 1.4), ("g53", 5.0 / 3.0)):
            f = np.array([oracle.riemann(kind, ql[i], qr[i], gamma, use_ref=True) for i in range(N)])
            out[f"riemann{kind}_{tag}_ql"], out[f"riemann{kind}_{tag}_qr"], out[f"riemann{kind}_{tag}_f"] = ql, qr, f
    vals = np.concatenate([rng.normal(0, 1, 60), [0.0, 0.0, 1.0, 1.0, -1.0, -1.0]])
    trip = np.stack([vals[:-2], vals[1:-1], vals[2:]], axis=1)
    nus = np.array([-0.7, -0.2, 0.0, 0.3, 0.9])
    out["slope_in"], out["slope_nu"] = trip, nus
    for st in range(9):
        out[f"slope{st}"] = np.array([[R.ref_compute_slope(a, b, c, st, 1.5, nu) for nu in nus] for a, b, c in trip])
    d = np.concatenate([rng.lognormal(0, 4, 60), [-1.0, 0.0, 1e-12, 1e-10]])
    p = np.concatenate([rng.lognormal(0, 4, 60), [1.0, 1e-30, 0.0, -2.0]])
    out["cs2_d"], out["cs2_p"], out["cs2_gamma"] = d, p, np.array([1.4, 5.0 / 3.0, 1.0001])
    out["cs2"] = np.array([[R.ref_get_cs2(d[i], p[i], g) for i in range(d.size)] for g in out["cs2_gamma"]])
    path = os.path.join(ROOT, "tests", "golden", "ref_kernel_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
