"""The oracle's composite godunov_fine pinned by checks that the mathematics
and the reference's own compiled sources fix, not by logs of a run.

What a composite check has to pin (HydroSolver.cpp:18-573 as the reference
builds its meshes, every face taking the coarse-ghost branch):
  * which states meet at each face: side 0 takes the neighbour's ghost qm
    and the cell's own qp, side 1 the cell's own qm and the neighbour's
    ghost qp (:447-452), both rotated into the face's frame (:455) and the
    flux rotated back (:475);
  * the 1/8 reflux of every face flux into the neighbour,
    unew(n) -= ((dtdx*sign)*F)*0.125 with sign +1 on side 0 (:476-482);
  * the cell's own update unew += dtdx*sum(sign*F) (:487-489) followed by
    its clamps (:534-547), cell by cell in processing order (level list
    order x ic, :565-573), so refluxes from cells processed later land
    after a cell's clamp;
  * the Courant step of the updated state (:878-977).

Checks:
  (a) brute force.  With slope_type = 0 and interpol_type = 0 every trace
      reduces to the identity (pinned by test_trace_zero_slope_is_identity),
      so each face flux is the Riemann solution of the two cells'
      primitive states.  `brute_godunov` recomputes the whole level from
      those: neighbours from cell coordinates (periodic), processing order
      from the level list, fluxes from the reference's own Riemann solver
      compiled from its sources (oracle/_ref; the oracle's solver, pinned
      bit for bit against it and the golden vectors, where _ref is not
      built), the reflux, the update and the clamps as the reference
      states them.  The oracle must agree value for value, including on a
      state where clamps fire before later refluxes land.  With slopes on,
      `brute_godunov` composes the pinned leaves (the reference's own
      compute_slope, the oracle's ctoprim and trace, the textbook ghost
      limiter) into own and ghost face states the same way.
  (b) symmetry.  For a state that varies along x only, the y and z momenta
      stay exactly zero (the two y faces of a cell carry the same flux and
      the two y refluxes into it cancel exactly) and every cell equals the
      cells with the same x up to rounding.
  (c) conservation.  With no clamp firing, the level's total change is
      (7/8)*dtdx*sum over cells of each cell's own flux sum: each cell
      keeps its flux sum and hands 1/8 of it to its neighbours.
  (d) Courant.  The two-step time-step sequence equals the closed-form
      candidate minimum of the initial state and of the brute-force state
      after one step.

test_planted_mistakes_fail builds the oracle with each planted mistake of
rh_oracle.cpp (reflux dropped, sign flipped, refluxes before the clamps,
factor 1/4, face states swapped) and shows each fails at least one check.
"""
import math

import numpy as np
import pytest

import oracle
import rhydro
from amr_meshes import cell_centres

GAMMA = 1.4


# ---------------------------------------------------------------- restatement

def _rot(q, d):
    q = np.array(q, dtype=np.float64)
    if d > 0:
        q[1], q[1 + d] = q[1 + d], q[1]
    return q


def _riemann_fn(kind):
    R = oracle.ref()
    if R is not None:
        return lambda ql, qr: oracle.riemann(kind, ql, qr, GAMMA, use_ref=True)
    return lambda ql, qr: oracle.riemann(kind, ql, qr, GAMMA)


def _sclamp(v, lo, hi):  # std::clamp: v < lo ? lo : (hi < v ? hi : v)
    return lo if v < lo else (hi if hi < v else v)


def _clamp_cell(u, gamma):
    """HydroSolver.cpp:534-547 on one cell's five conservative values."""
    d = _sclamp(u[0], 1e-10, 1e6)
    u[0] = d
    for i in (1, 2, 3):
        u[i] = _sclamp(u[i], -d * 10.0, d * 10.0)
    v2 = 0.0
    for i in (1, 2, 3):
        v = u[i] / d
        v2 = v2 + v * v
    ei = u[4] - 0.5 * d * v2
    floor = d * 1e-10 / (gamma - 1.0)
    if ei < floor:
        u[4] = floor + 0.5 * d * v2 + 0.0
    cap = d * 1e8
    u[4] = cap if cap < u[4] else u[4]


def _ghost_slope(dl, dr, it):
    """The ghost-state limiter (HydroSolver.cpp:415-431)."""
    if it > 0 and dl * dr > 0.0:
        sgn = 1.0 if dl >= 0.0 else -1.0
        if it == 1:
            return sgn * 0.5 * min(abs(dl), abs(dr))
        if it == 2:
            return sgn * min(abs(dl), abs(dr), 0.25 * abs(dl + dr))
        if it == 3:
            return 0.25 * (dl + dr)
    return 0.0


def _slope(ql, qc, qr, slope_type, nu):
    R = oracle.ref()
    if R is not None:
        return R.ref_compute_slope(ql, qc, qr, slope_type, 1.5, nu)
    return oracle.lib().or_compute_slope(ql, qc, qr, slope_type, 1.5, nu)


class Level:
    """A uniform periodic level in processing order, neighbours from cell
    coordinates (independent of the mesh's link arrays)."""

    def __init__(self, m, level):
        cells, cen = cell_centres(m, level)
        n = m.nx * 2 ** (level - 1)
        ijk = np.floor(cen * n / m.nx + 1e-9).astype(np.int64) % n
        where = {tuple(v): k for k, v in enumerate(ijk)}
        self.cells = cells
        self.ijk = ijk
        self.n = n
        self.nb = np.zeros((len(cells), 6), np.int64)
        for k, v in enumerate(ijk):
            for d in range(3):
                for s in range(2):
                    w = v.copy()
                    w[d] = (w[d] + (1 if s else -1)) % n
                    self.nb[k, 2 * d + s] = where[tuple(w)]


def brute_godunov(lv, uold, unew_in, dt, dx, kind, slope_type=0, interpol_type=0, gamma=GAMMA):
    """godunov_fine of one uniform periodic level from the reference's text
    and pinned leaves.  uold, unew_in: [5, ncells] in processing order.
    Returns (unew, flux_sums, clamped) with flux_sums[k] the cell's own
    sum(sign*F) and clamped[k] True when a clamp changed the cell."""
    riem = _riemann_fn(kind)
    N = uold.shape[1]
    q = np.array([oracle.ctoprim(uold[:, k], gamma) for k in range(N)])
    dtdx = dt / dx
    own = {}
    ghost = {}

    def own_faces(k, d):  # (qm, qp) of cell k along d, unrotated
        if (k, d) not in own:
            qc = q[k]
            if slope_type == 0:
                dq = np.zeros(5)
            else:
                ql, qr = q[lv.nb[k, 2 * d]], q[lv.nb[k, 2 * d + 1]]
                nu = qc[1 + d] * dt / dx
                dq = np.array([_slope(ql[i], qc[i], qr[i], slope_type, nu) / dx for i in range(5)])
            qm, qp = oracle.trace(_rot(qc, d), _rot(dq, d), dt, dx, gamma)
            own[(k, d)] = (_rot(qm, d), _rot(qp, d))
        return own[(k, d)]

    def ghost_faces(k, d):  # (qm, qp) of neighbour cell k as a coarse ghost along d
        if (k, d) not in ghost:
            qc = _rot(q[k], d)
            ql = _rot(q[lv.nb[k, 2 * d]], d)
            qr = _rot(q[lv.nb[k, 2 * d + 1]], d)
            dq = np.array([_ghost_slope(qc[i] - ql[i], qr[i] - qc[i], interpol_type) / (2.0 * dx)
                           for i in range(5)])
            qm, qp = oracle.trace(qc, dq, dt, 2.0 * dx, gamma)
            ghost[(k, d)] = (_rot(qm, d), _rot(qp, d))
        return ghost[(k, d)]

    unew = unew_in.copy()
    fsums = np.zeros((N, 5))
    clamped = np.zeros(N, bool)
    for k in range(N):
        fs = np.zeros(5)
        for d in range(3):
            for side in (0, 1):
                n = lv.nb[k, 2 * d + side]
                if side == 0:
                    ql, qr = ghost_faces(n, d)[0], own_faces(k, d)[1]
                else:
                    ql, qr = own_faces(k, d)[0], ghost_faces(n, d)[1]
                F = _rot(riem(_rot(ql, d), _rot(qr, d)), d)
                sign = 1.0 if side == 0 else -1.0
                unew[:, n] = unew[:, n] - ((dtdx * sign) * F) * 0.125
                fs = fs + sign * F
        fsums[k] = fs
        u = [float(v) for v in unew[:, k] + dtdx * fs]
        before = list(u)
        _clamp_cell(u, gamma)
        clamped[k] = any(a != b for a, b in zip(u, before))
        unew[:, k] = u
    return unew, fsums, clamped


def courant_closed_form(u, dx, gamma, cf, boxlen=1.0):
    """Candidate minimum of HydroSolver.cpp:931-973 without gravity
    (alpha = max(0, 1e-4) = 1e-4)."""
    dt = boxlen / 1e-10
    for k in range(u.shape[1]):
        d = max(u[0, k], 1e-10)
        v2 = 0.0
        for i in (1, 2, 3):
            v = u[i, k] / d
            v2 = v2 + v * v
        p = max((u[4, k] - 0.5 * d * v2 - 0.0) * (gamma - 1.0), d * 1e-10)
        p = min(p, d * 1e2)
        cs2 = oracle.lib().or_get_cs2(d, p, gamma) + 0.0 / d
        cs = math.sqrt(max(cs2, 1e-10))
        uw = 3 * cs
        for i in (1, 2, 3):
            uw = uw + abs(u[i, k] / d)
        alpha = max(0.0 * dx / (uw * uw), 0.0001)
        cand = (dx / uw) * (math.sqrt(1.0 + 2.0 * cf * alpha) - 1.0) / alpha
        dt = min(dt, cand)
    return dt


# ---------------------------------------------------------------- states

def _mesh(level):
    return rhydro.uniform_mesh(level)


def _set_state(m, lv, rho, v, p, gamma=GAMMA):
    """rho, p: [N]; v: [3, N] (processing order of lv)."""
    st = m.state("uold")
    for k, c in enumerate(lv.cells):
        e = p[k] / (gamma - 1.0) + 0.5 * rho[k] * (v[0, k] ** 2 + v[1, k] ** 2 + v[2, k] ** 2)
        st[:, c - 1] = [rho[k], rho[k] * v[0, k], rho[k] * v[1, k], rho[k] * v[2, k], e]
    m.unew[:] = m.uold


def x_wave(lv, amp=0.3):
    x = (lv.ijk[:, 0] + 0.5) / lv.n
    rho = 1.0 + amp * np.sin(2 * np.pi * x)
    p = 1.0 + 0.5 * amp * np.cos(2 * np.pi * x)
    v = np.zeros((3, len(x)))
    v[0] = 0.4 * np.sin(4 * np.pi * x)
    return rho, v, p


def random_state(lv, seed):
    r = np.random.default_rng(seed)
    N = len(lv.cells)
    rho = np.exp(0.4 * r.standard_normal(N))
    p = np.exp(0.4 * r.standard_normal(N))
    v = 0.5 * r.standard_normal((3, N))
    return rho, v, p


def clamp_state(lv):
    """Strong divergence at x = 0.5 and a near-vacuum there: with a large dt
    the density and momentum clamps fire in cells whose neighbours are
    processed later."""
    x = (lv.ijk[:, 0] + 0.5) / lv.n
    rho = np.where(np.abs(x - 0.5) < 0.3, 1e-3, 1.0)
    p = np.where(np.abs(x - 0.5) < 0.3, 1e-3, 1.0)
    v = np.zeros((3, len(x)))
    v[0] = np.where(x < 0.5, -3.0, 3.0)
    v[1] = 0.25 * np.sin(2 * np.pi * (lv.ijk[:, 1] + 0.5) / lv.n)
    return rho, v, p


def _run_oracle(m, lv, level, dt, dx, kind, slope_type, interpol_type, L=None):
    om = oracle.OracleMesh(m, L=L)
    p = oracle.params(GAMMA, slope_type, kind, interpol_type, 0)
    om.godunov_fine(p, level, dt, dx)
    return m.state("unew")[:, lv.cells - 1].copy()


def _same(a, b):
    return bool(((a == b) | (np.isnan(a) & np.isnan(b))).all())


CASES = [
    # (name, level, state, dt, riemann, slope_type, interpol_type)
    ("one oct, x wave, HLLC", 2, "xwave", 0.02, 2, 0, 0),
    ("4^3 octs, x wave, HLLC", 4, "xwave", 0.004, 2, 0, 0),
    ("4^3 octs, x wave, HLL", 4, "xwave", 0.004, 1, 0, 0),
    ("4^3 octs, x wave, LLF", 4, "xwave", 0.004, 0, 0, 0),
    ("2^3 octs, random 3D, HLLC", 3, "random", 0.004, 2, 0, 0),
    ("4^3 octs, clamps fire, HLLC", 4, "clamp", 0.2, 2, 0, 0),
    ("2^3 octs, random 3D, HLLC, MinMod + minmod ghosts", 3, "random", 0.004, 2, 1, 1),
    ("4^3 octs, x wave, HLLC, MC + central ghosts", 4, "xwave", 0.004, 2, 2, 3),
]


def _prepare(level, state):
    m = _mesh(level)
    lv = Level(m, level)
    rho, v, p = {"xwave": lambda: x_wave(lv), "random": lambda: random_state(lv, 11),
                 "clamp": lambda: clamp_state(lv)}[state]()
    _set_state(m, lv, rho, v, p)
    return m, lv


def _check_case(case, L=None):
    """True when the oracle (or the planted library L) equals the brute force."""
    _, level, state, dt, kind, st, it = case
    m, lv = _prepare(level, state)
    dx = 1.0 / 2 ** level
    uold = m.state("uold")[:, lv.cells - 1].copy()
    want, _, clamped = brute_godunov(lv, uold, uold.copy(), dt, dx, kind, st, it)
    got = _run_oracle(m, lv, level, dt, dx, kind, st, it, L)
    return _same(got, want), clamped


# ---------------------------------------------------------------- (a)

def test_zero_slope_traces_are_identities():
    q = np.array([1.3, -0.2, 0.7, 0.1, 2.5])
    qm, qp = oracle.trace(q, np.zeros(5), 0.01, 1 / 16, GAMMA)
    assert (qm == q).all() and (qp == q).all()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_brute_force_composite(case):
    ok, clamped = _check_case(case)
    assert ok, f"oracle godunov_fine differs from the brute force on {case[0]}"


def test_clamp_case_has_refluxes_landing_after_clamps():
    """In the clamp case some cell is clamped at its turn and then changed by
    a later neighbour's reflux, so the order of the two matters."""
    case = [c for c in CASES if c[2] == "clamp"][0]
    _, level, state, dt, kind, st, it = case
    m, lv = _prepare(level, state)
    dx = 1.0 / 2 ** level
    uold = m.state("uold")[:, lv.cells - 1].copy()
    unew, _, clamped = brute_godunov(lv, uold, uold.copy(), dt, dx, kind, st, it)
    assert clamped.sum() >= 8
    # A clamped cell whose final density is not a clamp bound: a reflux came
    # after its clamp.
    later = [k for k in np.flatnonzero(clamped)
             if any(lv.nb[k, f] > k for f in range(6)) and unew[0, k] not in (1e-10, 1e6)]
    assert later, "no clamped cell received a later reflux"


# ---------------------------------------------------------------- (b)

def test_x_only_state_symmetry():
    level = 4
    m, lv = _prepare(level, "xwave")
    dx = 1.0 / 2 ** level
    got = _run_oracle(m, lv, level, 0.004, dx, 2, 1, 1)
    # No clamp fires on this state (checked through the brute force).
    uold = m.state("uold")[:, lv.cells - 1].copy()
    want, _, clamped = brute_godunov(lv, uold, uold.copy(), 0.004, dx, 2, 1, 1)
    assert not clamped.any()
    assert (got[2] == 0.0).all() and (got[3] == 0.0).all(), "transverse momenta must stay exactly zero"
    for i in range(lv.n):
        sel = lv.ijk[:, 0] == i
        for iv in (0, 1, 4):
            col = got[iv, sel]
            assert np.allclose(col, col[0], rtol=4e-15, atol=0.0), f"x={i} var {iv} differs across y/z"


# ---------------------------------------------------------------- (c)

@pytest.mark.parametrize("state,st,it", [("random", 1, 1), ("xwave", 2, 3)])
def test_conservation_seven_eighths(state, st, it):
    level = 3
    m, lv = _prepare(level, state)
    dx, dt = 1.0 / 2 ** level, 0.004
    uold = m.state("uold")[:, lv.cells - 1].copy()
    want, fsums, clamped = brute_godunov(lv, uold, uold.copy(), dt, dx, 2, st, it)
    assert not clamped.any()
    got = _run_oracle(m, lv, level, dt, dx, 2, st, it)
    total = (got - uold).sum(axis=1)
    expect = 0.875 * (dt / dx) * fsums.sum(axis=0)
    scale = (dt / dx) * np.abs(fsums).sum(axis=0) + 1e-300
    assert (np.abs(total - expect) <= 1e-12 * scale).all(), (total, expect)
    # The relation is informative: the own flux sums do not cancel here.
    assert (np.abs(fsums.sum(axis=0)) > 1e-6 * np.abs(fsums).sum(axis=0))[[0, 4]].all()


# ---------------------------------------------------------------- (d)

def test_two_step_courant_sequence():
    level = 4
    m, lv = _prepare(level, "xwave")
    dx = 1.0 / 2 ** level
    uold = m.state("uold")[:, lv.cells - 1].copy()
    dt1 = courant_closed_form(uold, dx, GAMMA, 0.8)
    u1, _, _ = brute_godunov(lv, uold, uold.copy(), dt1, dx, 2)
    dt2 = courant_closed_form(u1, dx, GAMMA, 0.8)
    om = oracle.OracleMesh(m)
    dts = om.run_uniform(oracle.params(GAMMA, 0, 2, 0, 0), level, dx, 0.8, 2)
    assert dts[0] == dt1 and dts[1] == dt2
    assert dt1 != dt2


# ---------------------------------------------------------------- plants

PLANTS = {1: "reflux dropped", 2: "reflux sign flipped", 3: "refluxes before the clamps",
          4: "reflux factor 1/4", 5: "face states swapped"}


@pytest.mark.parametrize("plant", sorted(PLANTS), ids=[PLANTS[k] for k in sorted(PLANTS)])
def test_planted_mistakes_fail(plant, tmp_path):
    L = oracle.planted(plant, str(tmp_path))
    failed = [c[0] for c in CASES if not _check_case(c, L)[0]]
    assert failed, f"planted mistake '{PLANTS[plant]}' passes every brute-force case"
