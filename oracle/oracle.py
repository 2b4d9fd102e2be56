"""ctypes binding of the CPU oracle (liboracle.so) and of the reference's own
physics sources compiled in place (_ref/libref_kernels.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_kernels.so")

_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)


class OrMesh(C.Structure):
    _fields_ = [
        ("ncoarse", C.c_int), ("ngridmax", C.c_int), ("ncell", C.c_int), ("nvar", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("nlevelmax", C.c_int), ("ncpu", C.c_int), ("nboundary", C.c_int), ("myid", C.c_int),
        ("father", _IP), ("son", _IP), ("nbor", _IP), ("next", _IP), ("headl", _IP),
        ("uold", _DP), ("unew", _DP), ("f", _DP),
        ("bound_type", _IP), ("nbound_type", C.c_int),
    ]


class OrParams(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("slope_type", C.c_int), ("riemann", C.c_int),
        ("interpol_type", C.c_int), ("interpol_var", C.c_int), ("boxlen", C.c_double),
    ]


def build(quiet: bool = True) -> None:
    """Compiles liboracle.so (and _ref when the reference tree is present)."""
    out = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)


_lib = None
_ref = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _configure(C.CDLL(LIB_PATH))
    return _lib


def planted(plant: int, outdir: str) -> C.CDLL:
    """The oracle compiled with planted mistake `plant` (RH_ORACLE_PLANT in
    rh_oracle.cpp) into outdir, for tests that must show a check fails on a
    wrong restatement.  Same flags as the Makefile."""
    path = os.path.join(outdir, f"liboracle_plant{plant}.so")
    srcs = [os.path.join(HERE, f) for f in ("rh_oracle.cpp", "rh_oracle_amr.cpp", "rh_oracle_snapshot.cpp")]
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                    f"-DRH_ORACLE_PLANT={int(plant)}", "-o", path, *srcs], check=True, capture_output=True)
    return _configure(C.CDLL(path))


def _configure(L: C.CDLL) -> C.CDLL:
    L.or_godunov_fine.argtypes = [C.POINTER(OrMesh), C.POINTER(OrParams), C.c_int, C.c_double, C.c_double]
    L.or_godunov_fine_sample.argtypes = [C.POINTER(OrMesh), C.POINTER(OrParams), C.c_int, C.c_double,
                                         C.c_double, C.c_longlong]
    L.or_set_unew.argtypes = [C.POINTER(OrMesh), C.c_int]
    L.or_set_uold.argtypes = [C.POINTER(OrMesh), C.c_int]
    L.or_courant.argtypes = [C.POINTER(OrMesh), C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    L.or_courant.restype = C.c_double
    L.or_add_gravity_source_terms.argtypes = [C.POINTER(OrMesh), C.c_int, C.c_double]
    L.or_synchro_hydro_fine.argtypes = [C.POINTER(OrMesh), C.c_int, C.c_double, C.c_double]
    L.or_interpol_hydro.argtypes = [C.POINTER(OrParams), C.c_int, _DP, _DP]
    L.or_ctoprim.argtypes = [_DP, _DP, C.c_double]
    L.or_trace.argtypes = [_DP, _DP, C.c_double, C.c_double, _DP, _DP, C.c_double]
    L.or_compute_slope.argtypes = [C.c_double] * 3 + [C.c_int, C.c_double, C.c_double]
    L.or_compute_slope.restype = C.c_double
    L.or_ghost_slope.argtypes = [C.c_double, C.c_double, C.c_int]
    L.or_ghost_slope.restype = C.c_double
    L.or_get_cs2.argtypes = [C.c_double] * 3
    L.or_get_cs2.restype = C.c_double
    L.or_riemann.argtypes = [C.c_int, _DP, _DP, _DP, C.c_double]
    L.or_prim_to_cons.argtypes = [_DP, _DP, C.c_double]
    L.or_compute_flux.argtypes = [_DP, _DP, C.c_double]
    L.or_build_uniform.argtypes = [C.c_int] * 6 + [_IP] * 8
    L.or_refine.argtypes = [C.POINTER(OrMesh), C.POINTER(OrParams), C.c_int, _IP, C.c_int, _IP, _IP, _IP,
                            _DP, C.POINTER(C.c_int), _IP, _IP, _IP, _IP, _IP]
    return L


def ref() -> C.CDLL | None:
    """The reference's Riemann/slope/EOS sources (None when not built)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        R = C.CDLL(REF_PATH)
        R.ref_riemann.argtypes = [C.c_int, _DP, _DP, _DP, C.c_double]
        R.ref_compute_slope.argtypes = [C.c_double] * 3 + [C.c_int, C.c_double, C.c_double]
        R.ref_compute_slope.restype = C.c_double
        R.ref_get_cs2.argtypes = [C.c_double] * 3
        R.ref_get_cs2.restype = C.c_double
        _ref = R
    return _ref


def _ip(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_IP)


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def params(gamma=1.4, slope_type=1, riemann=0, interpol_type=1, interpol_var=0, boxlen=1.0) -> OrParams:
    return OrParams(gamma, slope_type, riemann, interpol_type, interpol_var, boxlen)


class OracleMesh:
    """Oracle view over AmrGrid-shaped numpy arrays (any object exposing
    father/son/nbor/next/headl/uold/unew and the scalars)."""

    def __init__(self, m, f: np.ndarray | None = None, L: C.CDLL | None = None):
        self.m = m
        self._L = L
        self._bt = np.zeros(1, np.int32)
        s = OrMesh()
        s.ncoarse, s.ngridmax, s.ncell, s.nvar = m.ncoarse, m.ngridmax, m.ncell, m.nvar
        s.nx, s.ny, s.nz = m.nx, m.ny, m.nz
        s.nlevelmax, s.ncpu, s.nboundary, s.myid = m.nlevelmax, 1, 0, 1
        s.father, s.son, s.nbor, s.next, s.headl = (_ip(getattr(m, k)) for k in
                                                    ("father", "son", "nbor", "next", "headl"))
        s.uold, s.unew = _dp(m.uold), _dp(m.unew)
        self._f = f
        s.f = _dp(f) if f is not None else C.cast(None, _DP)
        s.bound_type, s.nbound_type = _ip(self._bt), 0
        self.s = s

    def lib(self) -> C.CDLL:
        return self._L if self._L is not None else lib()

    def godunov_fine(self, p: OrParams, ilevel, dt, dx):
        st = self.lib().or_godunov_fine(C.byref(self.s), C.byref(p), ilevel, dt, dx)
        if st != 0:
            raise RuntimeError(f"oracle godunov_fine status {st}")

    def godunov_fine_sample(self, p: OrParams, ilevel, dt, dx, max_octs):
        st = self.lib().or_godunov_fine_sample(C.byref(self.s), C.byref(p), ilevel, dt, dx, max_octs)
        if st != 0:
            raise RuntimeError(f"oracle godunov_fine status {st}")

    def set_unew(self, ilevel):
        self.lib().or_set_unew(C.byref(self.s), ilevel)

    def set_uold(self, ilevel):
        self.lib().or_set_uold(C.byref(self.s), ilevel)

    def add_gravity_source_terms(self, ilevel, dt):
        self.lib().or_add_gravity_source_terms(C.byref(self.s), ilevel, dt)

    def synchro_hydro_fine(self, ilevel, dt, gamma):
        self.lib().or_synchro_hydro_fine(C.byref(self.s), ilevel, dt, gamma)

    def courant(self, ilevel, dx, gamma, courant_factor, boxlen=1.0) -> float:
        return self.lib().or_courant(C.byref(self.s), ilevel, dx, gamma, courant_factor, boxlen)

    def run_uniform(self, p: OrParams, ilevel, dx, courant_factor, nsteps) -> np.ndarray:
        """Simulation::amr_step(ilevel, 1) with levelmin == levelmax
        (Simulation.cpp:522-635): courant, set_unew, godunov_fine, set_uold."""
        dts = []
        for _ in range(nsteps):
            dt = self.courant(ilevel, dx, p.gamma, courant_factor, p.boxlen)
            self.set_unew(ilevel)
            self.godunov_fine(p, ilevel, dt, dx)
            self.set_uold(ilevel)
            dts.append(dt)
        return np.array(dts)


def interpol(p: OrParams, u1: np.ndarray, nvar: int = 5) -> np.ndarray:
    u1 = np.ascontiguousarray(u1, np.float64).reshape(7, 64)
    u2 = np.zeros((8, 64), np.float64)
    lib().or_interpol_hydro(C.byref(p), nvar, _dp(u1), _dp(u2))
    return u2


def riemann(kind: int, ql, qr, gamma: float, use_ref: bool = False) -> np.ndarray:
    ql = np.ascontiguousarray(ql, np.float64)
    qr = np.ascontiguousarray(qr, np.float64)
    f = np.zeros(5, np.float64)
    if use_ref:
        ref().ref_riemann(kind, _dp(ql), _dp(qr), _dp(f), gamma)
    else:
        lib().or_riemann(kind, _dp(ql), _dp(qr), _dp(f), gamma)
    return f


def ctoprim(u, gamma):
    u = np.ascontiguousarray(u, np.float64)
    q = np.zeros(5)
    lib().or_ctoprim(_dp(u), _dp(q), gamma)
    return q


def trace(q, dq, dt, dx, gamma):
    q = np.ascontiguousarray(q, np.float64)
    dq = np.ascontiguousarray(dq, np.float64)
    qm, qp = np.zeros(5), np.zeros(5)
    lib().or_trace(_dp(q), _dp(dq), dt, dx, _dp(qm), _dp(qp), gamma)
    return qm, qp


def build_uniform(nx, ny, nz, levelmin, nlevelmax, ngridmax):
    """The reference's own construction sequence (see rh_oracle.cpp)."""
    ncoarse = nx * ny * nz
    ncell = ncoarse + 8 * ngridmax
    a = dict(father=np.zeros(ngridmax, np.int32), son=np.zeros(ncell, np.int32),
             nbor=np.zeros(6 * ngridmax, np.int32), next=np.zeros(ngridmax, np.int32),
             prev=np.zeros(ngridmax, np.int32), headl=np.zeros(nlevelmax + 1, np.int32),
             taill=np.zeros(nlevelmax + 1, np.int32), numbl=np.zeros(nlevelmax + 1, np.int32))
    used = lib().or_build_uniform(nx, ny, nz, levelmin, nlevelmax, ngridmax,
                                  *[_ip(a[k]) for k in ("father", "son", "nbor", "next", "prev", "headl", "taill",
                                                        "numbl")])
    if used < 0:
        raise RuntimeError("ngridmax too small")
    return a, used


def refine(m, p: OrParams, ilevel: int, cells) -> int:
    """Refines the given 1-based cells of level list `ilevel` the way
    TreeUpdater::make_grid_fine does (see or_refine); `m` is a mesh object
    with the AmrGrid arrays, a `headf` attribute (first free grid) and an
    `epoch` that is bumped."""
    cells = np.ascontiguousarray(np.asarray(cells, dtype=np.int32))
    om = OracleMesh(m)
    headf = C.c_int(int(m.headf))
    n = lib().or_refine(C.byref(om.s), C.byref(p), ilevel, _ip(cells), int(cells.size), _ip(m.prev), _ip(m.taill),
                        _ip(m.numbl), _dp(m.xg), C.byref(headf), _ip(m.son), _ip(m.father), _ip(m.nbor),
                        _ip(m.next), _ip(m.headl))
    if n == -2:
        raise ValueError(f"level list {ilevel + 1} is beyond nlevelmax={m.nlevelmax}")
    if n < 0:
        raise RuntimeError("free list exhausted")
    m.headf = headf.value
    m.epoch += 1
    return n


# ---------------------------------------------------------------------------
# Adaptive run (rh_oracle_amr.cpp): the reference's tree updater and
# subcycled time integration around the oracle's hydro path.

class OrAmrConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ngridmax", C.c_int),
        ("levelmin", C.c_int), ("levelmax", C.c_int),
        ("boxlen", C.c_double), ("gamma", C.c_double), ("courant_factor", C.c_double),
        ("err_grad_d", C.c_double), ("err_grad_p", C.c_double), ("err_grad_v", C.c_double),
        ("nexpand", C.c_int), ("hydro", OrParams), ("ic_kind", C.c_int), ("nreg", C.c_int),
        ("reg_type", _IP),
        ("x_center", _DP), ("y_center", _DP), ("z_center", _DP),
        ("length_x", _DP), ("length_y", _DP), ("length_z", _DP),
        ("d_region", _DP), ("u_region", _DP), ("v_region", _DP), ("w_region", _DP), ("p_region", _DP),
        ("kh_amp", C.c_double),
    ]


class OrAmrState(C.Structure):
    _fields_ = [
        ("ncoarse", C.c_int), ("ngridmax", C.c_int), ("ncell", C.c_int), ("nlevelmax", C.c_int),
        ("father", _IP), ("son", _IP), ("nbor", _IP), ("next", _IP), ("headl", _IP), ("numbl", _IP),
        ("flag1", _IP), ("xg", _DP), ("uold", _DP), ("unew", _DP),
        ("t", C.c_double), ("nstep", C.c_int), ("dtnew", _DP), ("cell_updates", C.c_longlong),
        ("status", C.c_int),
    ]


def _amr_lib():
    L = lib()
    if not getattr(L, "_amr_ready", False):
        L.or_amr_create.argtypes = [C.POINTER(OrAmrConfig)]
        L.or_amr_create.restype = C.c_void_p
        L.or_amr_step.argtypes = [C.c_void_p, C.c_int]
        L.or_amr_get.argtypes = [C.c_void_p, C.POINTER(OrAmrState)]
        L.or_amr_destroy.argtypes = [C.c_void_p]
        L._amr_ready = True
    return L


class OracleAmr:
    """An adaptive run of the oracle: initialize on creation, then coarse
    steps.  `regions` as rhydro.init_regions takes them; ic='kh' selects the
    Kelvin-Helmholtz slab."""

    def __init__(self, nx=1, levelmin=4, levelmax=6, ngridmax=20000, gamma=1.4, courant_factor=0.8,
                 err_grad_d=-1.0, err_grad_p=-1.0, err_grad_v=-1.0, nexpand=1, slope_type=1, riemann=2,
                 interpol_type=1, interpol_var=0, boxlen=1.0, regions=None, ic="regions", kh_amp=0.01):
        regions = regions or []
        self._keep = []

        def col(key, default):
            a = np.array([float(r.get(key, default)) for r in regions] or [0.0], np.float64)
            self._keep.append(a)
            return a.ctypes.data_as(_DP)

        types = np.array([1 if r.get("type", "cube") == "sphere" else 0 for r in regions] or [0], np.int32)
        self._keep.append(types)
        cfg = OrAmrConfig()
        cfg.nx = cfg.ny = cfg.nz = nx
        cfg.ngridmax, cfg.levelmin, cfg.levelmax = ngridmax, levelmin, levelmax
        cfg.boxlen, cfg.gamma, cfg.courant_factor = boxlen, gamma, courant_factor
        cfg.err_grad_d, cfg.err_grad_p, cfg.err_grad_v = err_grad_d, err_grad_p, err_grad_v
        cfg.nexpand = nexpand
        cfg.hydro = OrParams(gamma, slope_type, riemann, interpol_type, interpol_var, boxlen)
        cfg.ic_kind = 1 if ic == "kh" else 0
        cfg.nreg = len(regions)
        cfg.reg_type = types.ctypes.data_as(_IP)
        (cfg.x_center, cfg.y_center, cfg.z_center, cfg.length_x, cfg.length_y, cfg.length_z, cfg.d_region,
         cfg.u_region, cfg.v_region, cfg.w_region, cfg.p_region) = (
            col("x_center", 0.5), col("y_center", 0.5), col("z_center", 0.5), col("length_x", 1.0),
            col("length_y", 1.0), col("length_z", 1.0), col("d", 1.0), col("u", 0.0), col("v", 0.0),
            col("w", 0.0), col("p", 1.0))
        cfg.kh_amp = kh_amp
        self._h = _amr_lib().or_amr_create(C.byref(cfg))
        if not self._h:
            raise ValueError("bad adaptive-run configuration")

    def step(self, n=1) -> int:
        return _amr_lib().or_amr_step(self._h, n)

    def state(self) -> dict:
        s = OrAmrState()
        _amr_lib().or_amr_get(self._h, C.byref(s))
        ng, nc, nl = s.ngridmax, s.ncell, s.nlevelmax

        def ints(p, n):
            return np.ctypeslib.as_array(p, shape=(n,)).copy()

        def dbl(p, n):
            return np.ctypeslib.as_array(p, shape=(n,)).copy()

        return dict(ncoarse=s.ncoarse, ngridmax=ng, ncell=nc, nlevelmax=nl, father=ints(s.father, ng),
                    son=ints(s.son, nc), nbor=ints(s.nbor, 6 * ng), next=ints(s.next, ng),
                    headl=ints(s.headl, nl + 1), numbl=ints(s.numbl, nl + 1), flag1=ints(s.flag1, nc),
                    xg=dbl(s.xg, 3 * ng), uold=dbl(s.uold, 5 * nc), unew=dbl(s.unew, 5 * nc), t=s.t,
                    nstep=s.nstep, dtnew=dbl(s.dtnew, 33), cell_updates=s.cell_updates, status=s.status)

    def dump(self, root: str, iout: int, **info) -> None:
        """dump_snapshot of the run into <root>/output_<iout> (rh_oracle_snapshot.cpp)."""
        st = _snap_lib().or_amr_dump(self._h, C.byref(snapshot_info(**info)), iout, os.fsencode(root))
        if st != 0:
            raise OSError(f"oracle snapshot into {root} failed")

    def close(self):
        if self._h:
            _amr_lib().or_amr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Snapshot writer (rh_oracle_snapshot.cpp): RamsesWriter's files, restated.

class OrSnapshotInfo(C.Structure):
    _fields_ = [("t", C.c_double)] + [(k, C.c_int) for k in ("nstep", "nstep_coarse", "noutput", "iout", "nener")] + [
        (k, C.c_double) for k in ("einit", "mass_tot_0", "rho_tot", "omega_m", "omega_l", "omega_k", "omega_b",
                                  "h0", "aexp_ini", "boxlen_ini", "aexp", "hexp", "aexp_old", "epot_tot_int",
                                  "epot_tot_old", "mass_sph")] + [
        ("levelmin", C.c_int), ("levelmax", C.c_int)] + [
        (k, C.c_double) for k in ("boxlen", "units_length", "units_density", "units_time")]


class OrSnapshotTree(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("nx", "ny", "nz", "ncoarse", "ngridmax", "ncell", "nlevelmax", "ncpu",
                                       "nboundary")] + [
        ("gamma", C.c_double), ("headl", _IP), ("taill", _IP), ("numbl", _IP),
        ("headf", C.c_int), ("tailf", C.c_int), ("numbf", C.c_int)] + [
        (k, _IP) for k in ("next", "prev", "father", "nbor", "son", "flag1", "cpu_map")] + [
        ("xg", _DP), ("phi", _DP), ("f", _DP), ("npart", C.c_longlong)]


def snapshot_info(**kw) -> OrSnapshotInfo:
    """SnapshotInfo's defaults (RamsesWriter.hpp:24-31; aexp 1, units 1),
    overridden by keyword."""
    i = OrSnapshotInfo()
    i.aexp_ini = i.boxlen_ini = i.aexp = i.aexp_old = 1.0
    i.units_length = i.units_density = i.units_time = 1.0
    i.boxlen = 1.0
    for k, v in kw.items():
        setattr(i, k, v)
    return i


def _snap_lib():
    L = lib()
    if not getattr(L, "_snap_ready", False):
        L.or_dump_snapshot.argtypes = [C.POINTER(OrSnapshotTree), _DP, C.POINTER(OrSnapshotInfo), C.c_char_p]
        L.or_amr_dump.argtypes = [C.c_void_p, C.POINTER(OrSnapshotInfo), C.c_int, C.c_char_p]
        L._snap_ready = True
    return L


def dump_snapshot(arrays: dict, dims: dict, root: str, **info) -> None:
    """dump_snapshot from host arrays: `arrays` holds the AmrGrid arrays by
    name (headl, taill, numbl, next, prev, father, nbor, son, xg, uold; flag1,
    cpu_map, phi, f optional), `dims` nx, ny, nz, ngridmax, nlevelmax, gamma
    and optionally headf, tailf, numbf, nboundary, npart."""
    t = OrSnapshotTree()
    t.nx, t.ny, t.nz = dims["nx"], dims["ny"], dims["nz"]
    t.ncoarse = t.nx * t.ny * t.nz
    t.ngridmax, t.nlevelmax = dims["ngridmax"], dims["nlevelmax"]
    t.ncell = t.ncoarse + 8 * t.ngridmax
    t.ncpu, t.nboundary = 1, dims.get("nboundary", 0)
    t.gamma = dims["gamma"]
    t.headf, t.tailf, t.numbf = dims.get("headf", 0), dims.get("tailf", 0), dims.get("numbf", 0)
    t.npart = dims.get("npart", 0)
    keep = []
    for k in ("headl", "taill", "numbl", "next", "prev", "father", "nbor", "son", "flag1", "cpu_map"):
        a = arrays.get(k)
        if a is None:
            setattr(t, k, C.cast(None, _IP))
        else:
            a = np.ascontiguousarray(a, np.int32)
            keep.append(a)
            setattr(t, k, _ip(a))
    for k in ("xg", "phi", "f"):
        a = arrays.get(k)
        if a is None:
            setattr(t, k, C.cast(None, _DP))
        else:
            a = np.ascontiguousarray(a, np.float64)
            keep.append(a)
            setattr(t, k, _dp(a))
    u = np.ascontiguousarray(arrays["uold"], np.float64)
    if _snap_lib().or_dump_snapshot(C.byref(t), _dp(u), C.byref(snapshot_info(**info)), os.fsencode(root)) != 0:
        raise OSError(f"oracle snapshot into {root} failed")


# ---- FLOP count of godunov_fine (flopcount.cpp) ------------------------------

FLOP_KINDS = ("add", "mul", "div", "sqrt", "cmp", "pow")
_flops = None


def _flops_lib() -> C.CDLL:
    global _flops
    if _flops is None:
        path = os.path.join(HERE, "liboracle_flops.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_flop_count.argtypes = [C.c_int] * 6 + [_IP] * 5 + [_DP, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                                                 C.c_double, C.c_double, _DP,
                                                                 C.POINTER(C.c_ulonglong)]
        L.or_flop_count.restype = C.c_int
        _flops = L
    return _flops


def flop_count(m, ilevel: int, gamma: float, slope_type: int, riemann: int, interpol_type: int, dt: float,
               dx: float):
    """set_unew + godunov_fine of `ilevel` on a copy of mesh `m` through the
    oracle with a counting floating-point type.  Returns (unew after the
    call, {kind: count} for the godunov_fine call); `m` is not modified."""
    unew = np.zeros_like(m.uold)
    counts = (C.c_ulonglong * 6)()
    arrs = [np.ascontiguousarray(getattr(m, k), np.int32) for k in ("father", "son", "nbor", "next", "headl")]
    uold = np.ascontiguousarray(m.uold, np.float64)
    st = _flops_lib().or_flop_count(m.ncoarse, m.ngridmax, m.nx, m.ny, m.nz, m.nlevelmax, *(_ip(a) for a in arrs),
                                    _dp(uold), ilevel, gamma, slope_type, riemann, interpol_type, dt, dx,
                                    _dp(unew), counts)
    if st != 0:
        raise RuntimeError(f"counting oracle status {st}")
    return unew, dict(zip(FLOP_KINDS, (int(c) for c in counts)))
