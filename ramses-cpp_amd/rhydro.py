"""ctypes binding of librhydro.so (include/rhydro.h, include/rhydro_driver.h).

Python is plumbing here: it holds the AmrGrid-shaped numpy arrays of a run
and calls the C-ABI.  All hydro arithmetic happens in the HIP kernels of the
library.  Loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librhydro.so")

RH_OK, RH_ERR_ARG, RH_ERR_HIP, RH_ERR_UNSUPPORTED, RH_ERR_TOPOLOGY, RH_ERR_NO_DEVICE, RH_ERR_HOST = range(7)
RIEMANN = {"llf": 0, "hll": 1, "hllc": 2, "exact": 3, "acoustic": 4}
RH_PREC_FP64, RH_PREC_MIXED = 0, 1
RH_HOST_COHERENT, RH_DEVICE_RESIDENT = 0, 1
RH_SYNC_TO_HOST, RH_SYNC_TO_DEVICE = 1, 2
RH_UOLD, RH_UNEW = 1, 2
TK_GODUNOV, TK_COURANT, TK_SET_UNEW, TK_SET_UOLD, TK_STEP_JOIN = range(5)
RH_KERNEL_VARIANTS = 7  # 0 per-cell; 1, 2 oct-in-lanes with 16 / 8 octs per workgroup; 3, 4, 5 bricks of 16 / 32 / 64 octs; 6 one wave per dimension

_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)


class RhParams(C.Structure):
    _fields_ = [
        ("gamma", C.c_double),
        ("slope_type", C.c_int),
        ("riemann", C.c_int),
        ("interpol_type", C.c_int),
        ("interpol_var", C.c_int),
        ("nvar", C.c_int),
        ("nener", C.c_int),
        ("barotropic_eos", C.c_int),
        ("boxlen", C.c_double),
        ("precision", C.c_int),
        ("device", C.c_int),
    ]


class RhMesh(C.Structure):
    _fields_ = [
        ("ncoarse", C.c_int), ("ngridmax", C.c_int), ("ncell", C.c_int), ("nvar", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("nlevelmax", C.c_int), ("ncpu", C.c_int), ("nboundary", C.c_int), ("myid", C.c_int),
        ("father", _IP), ("son", _IP), ("nbor", _IP), ("next", _IP), ("headl", _IP),
        ("uold", _DP), ("unew", _DP), ("f", _DP),
        ("mesh_epoch", C.c_longlong),
    ]


class RhSnapshotInfo(C.Structure):
    """rh_snapshot_info (rhydro_snapshot.h): SnapshotInfo plus the namelist
    values the info file prints."""
    _fields_ = [("t", C.c_double)] + [(k, C.c_int) for k in ("nstep", "nstep_coarse", "noutput", "iout", "nener")] + [
        (k, C.c_double) for k in ("einit", "mass_tot_0", "rho_tot", "omega_m", "omega_l", "omega_k", "omega_b",
                                  "h0", "aexp_ini", "boxlen_ini", "aexp", "hexp", "aexp_old", "epot_tot_int",
                                  "epot_tot_old", "mass_sph")] + [
        ("levelmin", C.c_int), ("levelmax", C.c_int)] + [
        (k, C.c_double) for k in ("boxlen", "units_length", "units_density", "units_time")]


class RhSnapshotTree(C.Structure):
    """rh_snapshot_tree (rhydro_snapshot.h): the AmrGrid arrays the writers read."""
    _fields_ = [(k, C.c_int) for k in ("nx", "ny", "nz", "ncoarse", "ngridmax", "ncell", "nlevelmax", "ncpu",
                                       "nboundary")] + [
        ("gamma", C.c_double), ("headl", C.POINTER(C.c_int)), ("taill", C.POINTER(C.c_int)),
        ("numbl", C.POINTER(C.c_int)), ("headf", C.c_int), ("tailf", C.c_int), ("numbf", C.c_int)] + [
        (k, C.POINTER(C.c_int)) for k in ("next", "prev", "father", "nbor", "son", "flag1", "cpu_map")] + [
        ("xg", C.POINTER(C.c_double)), ("phi", C.POINTER(C.c_double)), ("f", C.POINTER(C.c_double)),
        ("npart", C.c_longlong)]


def snapshot_info(**kw) -> RhSnapshotInfo:
    """SnapshotInfo's defaults (RamsesWriter.hpp:24-31: aexp 1, no outputs)
    with unit scales 1 and boxlen 1, overridden by keyword."""
    i = RhSnapshotInfo()
    i.aexp_ini = i.boxlen_ini = i.aexp = i.aexp_old = 1.0
    i.units_length = i.units_density = i.units_time = 1.0
    i.boxlen = 1.0
    for k, v in kw.items():
        setattr(i, k, v)
    return i


def resolve_riemann(name: str | None) -> int:
    """Namelist string -> solver id, as the reference's face loop resolves it
    (HydroSolver.cpp:456-461): exact, case-sensitive matches, else LLF."""
    return RIEMANN.get(name, 0) if name in ("hllc", "hll", "exact", "acoustic") else 0


class RhAmrParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ngridmax", C.c_int),
        ("levelmin", C.c_int), ("levelmax", C.c_int), ("courant_factor", C.c_double),
        ("err_grad_d", C.c_double), ("err_grad_p", C.c_double), ("err_grad_v", C.c_double),
        ("nexpand", C.c_int), ("ic_kind", C.c_int), ("nreg", C.c_int), ("reg_type", C.POINTER(C.c_int)),
        ("x_center", C.POINTER(C.c_double)), ("y_center", C.POINTER(C.c_double)),
        ("z_center", C.POINTER(C.c_double)), ("length_x", C.POINTER(C.c_double)),
        ("length_y", C.POINTER(C.c_double)), ("length_z", C.POINTER(C.c_double)),
        ("d_region", C.POINTER(C.c_double)), ("u_region", C.POINTER(C.c_double)),
        ("v_region", C.POINTER(C.c_double)), ("w_region", C.POINTER(C.c_double)),
        ("p_region", C.POINTER(C.c_double)), ("kh_amp", C.c_double),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.rh_create.argtypes = [C.POINTER(RhParams), C.POINTER(C.c_void_p)]
        L.rh_destroy.argtypes = [C.c_void_p]
        L.rh_bind.argtypes = [C.c_void_p, C.POINTER(RhMesh), C.c_int]
        L.rh_set_unew.argtypes = [C.c_void_p, C.c_int]
        L.rh_set_uold.argtypes = [C.c_void_p, C.c_int]
        L.rh_godunov_fine.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        L.rh_courant.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, _DP]
        L.rh_courant_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, _DP]
        L.rh_interpol_hydro.argtypes = [C.c_void_p, C.c_int, _DP, _DP]
        L.rh_sync.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.rh_last_error.argtypes = [C.c_void_p]
        L.rh_last_error.restype = C.c_char_p
        L.rh_run_uniform.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, _DP]
        L.rh_uniform_ngrid.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.rh_uniform_ngrid.restype = C.c_longlong
        L.rh_build_uniform.argtypes = [C.c_int] * 6 + [_IP] * 8 + [_DP]
        L.rh_build_uniform.restype = C.c_longlong
        L.rh_init_regions.argtypes = [C.POINTER(RhMesh), _DP, C.c_double, C.c_double, C.c_int, _IP] + [_DP] * 11
        L.rh_init_lognormal.argtypes = [C.POINTER(RhMesh), C.c_int, C.c_ulonglong, C.c_double, C.c_double,
                                        C.c_double]
        L.rh_timers_enable.argtypes = [C.c_void_p, C.c_int]
        L.rh_timers_read.argtypes = [C.c_void_p, C.c_int, _DP, C.POINTER(C.c_longlong)]
        L.rh_timers_reset.argtypes = [C.c_void_p]
        L.rh_level_leaf_cells.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_longlong)]
        L.rh_set_kernel_variant.argtypes = [C.c_void_p, C.c_int]
        L.rh_get_kernel_variant.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.rh_autotune.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_int), _DP]
        L.rh_set_graphs.argtypes = [C.c_void_p, C.c_int]
        L.rh_set_unew_oct_lanes.argtypes = [C.c_void_p, C.c_int]
        L.rh_graph_stats.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.rh_add_gravity_source_terms.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.rh_synchro_hydro_fine.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.rh_amr_create.argtypes = [C.POINTER(RhParams), C.POINTER(RhAmrParams), C.POINTER(C.c_void_p)]
        L.rh_amr_destroy.argtypes = [C.c_void_p]
        L.rh_amr_step.argtypes = [C.c_void_p, C.c_int]
        L.rh_amr_mesh.argtypes = [C.c_void_p, C.POINTER(RhMesh), C.POINTER(_IP), C.POINTER(_DP)]
        L.rh_amr_stats.argtypes = [C.c_void_p, _DP, _IP, C.POINTER(C.c_longlong)]
        L.rh_amr_context.argtypes = [C.c_void_p]
        L.rh_amr_context.restype = C.c_void_p
        L.rh_amr_timers.argtypes = [C.c_void_p, _DP]
        L.rh_init_shock_tube.argtypes = [C.POINTER(RhMesh), _DP, C.c_double, C.c_double, _DP, _DP, _DP, _DP]
        L.rh_bind_ensemble.argtypes = [C.c_void_p, C.c_int, C.POINTER(RhMesh), C.c_int]
        L.rh_run_ensemble.argtypes = [C.c_void_p, _IP, _DP, C.c_double, C.c_int, _DP]
        L.rh_ensemble_leaf_cells.argtypes = [C.c_void_p, _IP, C.POINTER(C.c_longlong)]
        L.rh_dump_snapshot.argtypes = [C.c_void_p, C.POINTER(RhSnapshotTree), C.POINTER(RhSnapshotInfo),
                                       C.c_char_p]
        L.rh_write_hydro.argtypes = [C.c_void_p, C.POINTER(RhSnapshotTree), C.POINTER(RhSnapshotInfo), C.c_char_p]
        L.rh_write_amr.argtypes = [C.POINTER(RhSnapshotTree), C.POINTER(RhSnapshotInfo), C.c_char_p]
        L.rh_amr_dump.argtypes = [C.c_void_p, C.POINTER(RhSnapshotInfo), C.c_int, C.c_char_p]
        _lib = L
    return _lib


EXPORTED = [
    "rh_create", "rh_destroy", "rh_bind", "rh_set_unew", "rh_godunov_fine", "rh_set_uold", "rh_courant",
    "rh_courant_noalloc",
    "rh_interpol_hydro", "rh_sync", "rh_last_error", "rh_run_uniform", "rh_uniform_ngrid", "rh_build_uniform",
    "rh_init_regions", "rh_init_lognormal", "rh_timers_enable", "rh_timers_read", "rh_timers_reset",
    "rh_level_leaf_cells", "rh_set_kernel_variant", "rh_get_kernel_variant", "rh_autotune", "rh_init_shock_tube",
    "rh_bind_ensemble", "rh_run_ensemble", "rh_ensemble_leaf_cells", "rh_add_gravity_source_terms",
    "rh_synchro_hydro_fine", "rh_amr_create", "rh_amr_destroy", "rh_amr_step", "rh_amr_mesh", "rh_amr_stats",
    "rh_amr_context", "rh_amr_timers", "rh_dump_snapshot", "rh_write_hydro", "rh_write_amr", "rh_amr_dump",
    "rh_set_graphs", "rh_graph_stats", "rh_set_unew_oct_lanes",
]


def _ip(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_IP)


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


@dataclass
class Mesh:
    """AmrGrid-shaped arrays (AmrGrid.hpp:15-33) for ncpu = 1."""
    nx: int
    ny: int
    nz: int
    ngridmax: int
    nlevelmax: int
    nvar: int = 5
    epoch: int = 0
    arrays: dict = field(default_factory=dict)

    def __post_init__(self):
        self.ncoarse = self.nx * self.ny * self.nz
        self.ncell = self.ncoarse + 8 * self.ngridmax
        a = self.arrays
        a["father"] = np.zeros(self.ngridmax, np.int32)
        a["son"] = np.zeros(self.ncell, np.int32)
        a["nbor"] = np.zeros(6 * self.ngridmax, np.int32)
        a["next"] = np.zeros(self.ngridmax, np.int32)
        a["prev"] = np.zeros(self.ngridmax, np.int32)
        a["headl"] = np.zeros(self.nlevelmax + 1, np.int32)
        a["taill"] = np.zeros(self.nlevelmax + 1, np.int32)
        a["numbl"] = np.zeros(self.nlevelmax + 1, np.int32)
        a["xg"] = np.zeros(3 * self.ngridmax, np.float64)
        a["uold"] = np.zeros(self.nvar * self.ncell, np.float64)
        a["unew"] = np.zeros(self.nvar * self.ncell, np.float64)

    def __getattr__(self, k):
        arrays = self.__dict__.get("arrays")
        if arrays is not None and k in arrays:
            return arrays[k]
        raise AttributeError(k)

    def rh_mesh(self, f: np.ndarray | None = None) -> RhMesh:
        a = self.arrays
        m = RhMesh()
        m.ncoarse, m.ngridmax, m.ncell, m.nvar = self.ncoarse, self.ngridmax, self.ncell, self.nvar
        m.nx, m.ny, m.nz = self.nx, self.ny, self.nz
        m.nlevelmax, m.ncpu, m.nboundary, m.myid = self.nlevelmax, 1, 0, 1
        m.father, m.son, m.nbor, m.next, m.headl = (_ip(a[k]) for k in ("father", "son", "nbor", "next", "headl"))
        m.uold, m.unew = _dp(a["uold"]), _dp(a["unew"])
        m.f = _dp(f) if f is not None else C.cast(None, _DP)
        m.mesh_epoch = self.epoch
        return m

    def level_octs(self, level: int) -> list[int]:
        out, g = [], int(self.headl[level])
        while g > 0:
            out.append(g)
            g = int(self.next[g - 1])
        return out

    def cells_of_level(self, level: int, leaf_only: bool = False) -> np.ndarray:
        """0-based cell indices of the level's octs in processing order."""
        octs = np.array(self.level_octs(level), dtype=np.int64)
        ic = np.arange(8, dtype=np.int64)
        cells = (self.ncoarse + ic[None, :] * self.ngridmax + octs[:, None] - 1).reshape(-1)
        if leaf_only:
            cells = cells[self.son[cells] == 0]
        return cells

    def state(self, which: str = "uold") -> np.ndarray:
        return self.arrays[which].reshape(self.nvar, self.ncell)

    def snapshot_tree(self, gamma: float, flag1=None, cpu_map=None, phi=None, f=None):
        """(rh_snapshot_tree, arrays it points to) for the writers; flag1 and
        cpu_map default to 0 and 1, phi and f to 0."""
        a = self.arrays
        t = RhSnapshotTree()
        t.nx, t.ny, t.nz, t.ncoarse, t.ngridmax, t.ncell = self.nx, self.ny, self.nz, self.ncoarse, self.ngridmax, \
            self.ncell
        t.nlevelmax, t.ncpu, t.nboundary, t.gamma = self.nlevelmax, 1, 0, gamma
        t.headf, t.tailf, t.numbf = (int(getattr(self, k, 0)) for k in ("headf", "tailf", "numbf"))
        keep = []
        for k in ("headl", "taill", "numbl", "next", "prev", "father", "nbor", "son"):
            setattr(t, k, _ip(a[k]))
        for k, v in (("flag1", flag1), ("cpu_map", cpu_map)):
            if v is None:
                setattr(t, k, C.cast(None, _IP))
            else:
                v = np.ascontiguousarray(v, np.int32)
                keep.append(v)
                setattr(t, k, _ip(v))
        t.xg = _dp(a["xg"])
        for k, v in (("phi", phi), ("f", f)):
            if v is None:
                setattr(t, k, C.cast(None, _DP))
            else:
                v = np.ascontiguousarray(v, np.float64)
                keep.append(v)
                setattr(t, k, _dp(v))
        t.npart = 0
        return t, keep


def uniform_mesh(levelmin: int, nx: int = 1, ny: int = 1, nz: int = 1, ngridmax: int | None = None,
                 nlevelmax: int | None = None) -> Mesh:
    """The uniform mesh the reference builds for levelmin == levelmax."""
    L = lib()
    need = int(L.rh_uniform_ngrid(nx, ny, nz, levelmin))
    ngridmax = max(need, 1) if ngridmax is None else ngridmax
    nlevelmax = levelmin if nlevelmax is None else nlevelmax
    m = Mesh(nx, ny, nz, ngridmax, nlevelmax)
    a = m.arrays
    ng = L.rh_build_uniform(nx, ny, nz, levelmin, nlevelmax, ngridmax, _ip(a["father"]), _ip(a["son"]),
                            _ip(a["nbor"]), _ip(a["next"]), _ip(a["prev"]), _ip(a["headl"]), _ip(a["taill"]),
                            _ip(a["numbl"]), _dp(a["xg"]))
    if ng < 0:
        raise RuntimeError(f"rh_build_uniform failed ({ng})")
    # The free list after the build: grids ng+1..ngridmax (AmrGrid.cpp:259-302).
    m.headf = int(ng) + 1 if ng < ngridmax else 0
    m.tailf = ngridmax if ng < ngridmax else 0
    m.numbf = ngridmax - int(ng)
    return m


def write_amr(mesh: Mesh, path: str, gamma: float = 1.4, flag1=None, cpu_map=None, **info) -> None:
    """RamsesWriter::write_amr of a host mesh (no context needed)."""
    t, keep = mesh.snapshot_tree(gamma, flag1, cpu_map)
    st = lib().rh_write_amr(C.byref(t), C.byref(snapshot_info(**info)), os.fsencode(path))
    if st != RH_OK:
        raise RhError(st, lib().rh_last_error(None).decode())


def init_regions(m: Mesh, regions: list[dict], gamma: float, boxlen: float = 1.0) -> None:
    """regions: dicts with type ('square'/'cube'/'sphere'), x/y/z_center,
    length_x/y/z, d, u, v, w, p (Initializer.cpp defaults for missing keys)."""
    n = len(regions)

    def col(key, default):
        return np.array([float(r.get(key, default)) for r in regions], np.float64)

    types = np.array([1 if r.get("type", "cube") == "sphere" else 0 for r in regions], np.int32)
    cols = [col("x_center", 0.5), col("y_center", 0.5), col("z_center", 0.5), col("length_x", 1.0),
            col("length_y", 1.0), col("length_z", 1.0), col("d", 1.0), col("u", 0.0), col("v", 0.0),
            col("w", 0.0), col("p", 1.0)]
    mm = m.rh_mesh()
    st = lib().rh_init_regions(C.byref(mm), _dp(m.xg), boxlen, gamma, n, _ip(types), *[_dp(c) for c in cols])
    if st != RH_OK:
        raise RuntimeError(f"rh_init_regions failed ({st})")


def init_lognormal(m: Mesh, level: int, seed: int, gamma: float, sigma: float = 0.5, vel_amp: float = 0.5) -> None:
    mm = m.rh_mesh()
    st = lib().rh_init_lognormal(C.byref(mm), level, seed, gamma, sigma, vel_amp)
    if st != RH_OK:
        raise RuntimeError(f"rh_init_lognormal failed ({st})")


def sedov_regions(levelmin: int, nx: int = 1, gamma: float = 1.4, energy: float = 1.0, p0: float = 1e-5,
                  boxlen: float = 1.0) -> list[dict]:
    """Sedov set-up of the benchmark: uniform rho=1, p=p0 square, and the
    energy deposited as pressure in the cells with r < 1.5 h (8 cells)."""
    h = boxlen / (nx * 2 ** (levelmin - 1))
    p_in = (gamma - 1.0) * energy / (8 * h ** 3)
    return [
        dict(type="square", x_center=0.5, y_center=0.5, z_center=0.5, length_x=10.0, length_y=10.0,
             length_z=10.0, d=1.0, p=p0),
        dict(type="sphere", x_center=0.5, y_center=0.5, z_center=0.5, length_x=1.5 * h, d=1.0, p=p_in),
    ]


def init_shock_tube(m: Mesh, normal, point, left, right, gamma: float, boxlen: float = 1.0) -> None:
    """Two primitive states (d, u, v, w, p) split by the plane through
    `point` with normal `normal` (rh_init_shock_tube)."""
    arrs = [np.ascontiguousarray(a, np.float64) for a in (normal, point, left, right)]
    mm = m.rh_mesh()
    st = lib().rh_init_shock_tube(C.byref(mm), _dp(m.xg), boxlen, gamma, *[_dp(a) for a in arrs])
    if st != RH_OK:
        raise RuntimeError(f"rh_init_shock_tube failed ({st})")


# Box sides of the shock-tube ensemble: nx * 2^(L-1) cells per dimension with
# nx coarse cells (nx in 1, 3, 5, 7) and reference level L >= 2.
def _quantised_sides(lo: int, hi: int) -> list[tuple[int, int, int]]:
    out = []
    for nx in (1, 3, 5, 7):
        for L in range(2, 16):
            side = nx * 2 ** (L - 1)
            if lo <= side <= hi:
                out.append((side, nx, L))
    return sorted(out)


@dataclass
class EnsembleBox:
    mesh: Mesh
    nx: int
    level: int
    dx: float
    kind: str = "sod"                 # "sod" (shock tube) or "sedov" (central energy)
    normal: np.ndarray | None = None  # sod: plane normal
    left: np.ndarray | None = None    # sod: primitive state on the normal's negative side
    right: np.ndarray | None = None
    energy: float = 0.0               # sedov: deposited energy


def box_ensemble(nbox: int = 96, seed: int = 20261202, side_min: int = 16, side_max: int = 64, gamma: float = 1.4,
                 boxlen: float = 1.0, sedov_share: float = 0.5) -> list[EnsembleBox]:
    """Independent periodic cubic boxes of the batched benchmark.  Sides are
    drawn log-uniformly in [side_min, side_max] from
    numpy.random.default_rng(seed) and rounded (in log) to the nearest
    nx * 2^(L-1).  With probability sedov_share a box is a Sedov blast: rho = 1,
    p = 1e-5 at rest, and an energy drawn uniformly in [0.1, 1] deposited as
    pressure in the 8 cells within 1.5 cell widths of the centre
    (Initializer.cpp:87-94, the set-up of sedov_regions); otherwise it is a
    shock tube through the centre, the plane normal uniform on the sphere and
    each side's density and pressure uniform in [0.1, 1], at rest.  The
    solver's cell width is dx = boxlen / (nx * 2^L), as the reference passes
    it (Simulation.cpp:548).  With sedov_share = 0 no kind is drawn, so the
    boxes are those of shock_tube_ensemble for the same seed."""
    rng = np.random.default_rng(seed)
    table = _quantised_sides(side_min, side_max)
    logs = np.log([t[0] for t in table])
    boxes = []
    for _ in range(nbox):
        s = rng.uniform(np.log(side_min), np.log(side_max))
        side, nx, L = table[int(np.argmin(np.abs(logs - s)))]
        m = uniform_mesh(L, nx=nx, ny=nx, nz=nx)
        dx = boxlen / (nx * 2 ** L)
        if sedov_share > 0 and rng.uniform() < sedov_share:
            e = float(rng.uniform(0.1, 1.0))
            init_regions(m, sedov_regions(L, nx=nx, gamma=gamma, energy=e, boxlen=boxlen), gamma, boxlen)
            boxes.append(EnsembleBox(m, nx, L, dx, kind="sedov", energy=e))
            continue
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        dl, pl, dr, pr = rng.uniform(0.1, 1.0, 4)
        left = np.array([dl, 0.0, 0.0, 0.0, pl])
        right = np.array([dr, 0.0, 0.0, 0.0, pr])
        init_shock_tube(m, n, np.full(3, 0.5 * boxlen), left, right, gamma, boxlen)
        boxes.append(EnsembleBox(m, nx, L, dx, kind="sod", normal=n, left=left, right=right))
    return boxes


def shock_tube_ensemble(nbox: int = 96, seed: int = 20261202, side_min: int = 16, side_max: int = 64,
                        gamma: float = 1.4, boxlen: float = 1.0) -> list[EnsembleBox]:
    """The shock-tube boxes of box_ensemble only (sedov_share = 0)."""
    return box_ensemble(nbox, seed, side_min, side_max, gamma, boxlen, sedov_share=0.0)


class RhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rhydro error {code}: {msg}")
        self.code = code


class Solver:
    """One rh_ctx bound to one Mesh (the HydroSolver of the reference)."""

    def __init__(self, gamma=1.4, slope_type=1, riemann="hllc", interpol_type=1, interpol_var=0, boxlen=1.0,
                 precision=RH_PREC_FP64, device=0, variant=None):
        self.params = RhParams(gamma, slope_type, riemann if isinstance(riemann, int) else resolve_riemann(riemann),
                               interpol_type, interpol_var, 5, 0, 0, boxlen, precision, device)
        self._ctx = C.c_void_p()
        st = lib().rh_create(C.byref(self.params), C.byref(self._ctx))
        if st != RH_OK:
            raise RhError(st, lib().rh_last_error(None).decode())
        self.mesh = None
        self._rhm = None
        self._f = None
        if variant is not None:
            self._chk(lib().rh_set_kernel_variant(self._ctx, variant))

    def _chk(self, st):
        if st != RH_OK:
            raise RhError(st, lib().rh_last_error(self._ctx).decode())

    def bind(self, mesh: Mesh, mode=RH_DEVICE_RESIDENT, f: np.ndarray | None = None):
        self.mesh, self._f = mesh, f
        self._rhm = mesh.rh_mesh(f)
        self._chk(lib().rh_bind(self._ctx, C.byref(self._rhm), mode))

    def bind_ensemble(self, meshes: list, mode=RH_DEVICE_RESIDENT):
        """Binds independent meshes as one ensemble (rh_bind_ensemble)."""
        self.mesh, self._f = list(meshes), None
        arr = (RhMesh * len(meshes))(*[m.rh_mesh() for m in meshes])
        self._rhm = arr
        self._chk(lib().rh_bind_ensemble(self._ctx, len(meshes), arr, mode))

    def run_ensemble(self, levels, dx, courant_factor, nsteps) -> np.ndarray:
        """Advances every box nsteps steps; returns the time steps, shape
        (nsteps, nbox)."""
        lv = np.ascontiguousarray(levels, np.int32)
        dxa = np.ascontiguousarray(dx, np.float64)
        nbox = len(lv)
        dts = np.zeros(max(nsteps, 1) * nbox, np.float64)
        self._chk(lib().rh_run_ensemble(self._ctx, _ip(lv), _dp(dxa), courant_factor, nsteps, _dp(dts)))
        return dts[: nsteps * nbox].reshape(nsteps, nbox)

    def ensemble_leaf_cells(self, levels) -> int:
        lv = np.ascontiguousarray(levels, np.int32)
        n = C.c_longlong()
        self._chk(lib().rh_ensemble_leaf_cells(self._ctx, _ip(lv), C.byref(n)))
        return n.value

    def set_unew(self, ilevel):
        self._chk(lib().rh_set_unew(self._ctx, ilevel))

    def set_uold(self, ilevel):
        self._chk(lib().rh_set_uold(self._ctx, ilevel))

    def godunov_fine(self, ilevel, dt, dx):
        self._chk(lib().rh_godunov_fine(self._ctx, ilevel, dt, dx))

    def add_gravity_source_terms(self, ilevel, dt):
        self._chk(lib().rh_add_gravity_source_terms(self._ctx, ilevel, dt))

    def synchro_hydro_fine(self, ilevel, dt):
        self._chk(lib().rh_synchro_hydro_fine(self._ctx, ilevel, dt))

    def dump_snapshot(self, mesh: Mesh, root: str, gamma: float | None = None, flag1=None, cpu_map=None, phi=None,
                      f=None, **info):
        """Simulation::dump_snapshot of the bound mesh into <root>/output_<iout>
        (hydro records packed on the GPU from the device state)."""
        t, keep = mesh.snapshot_tree(self.params.gamma if gamma is None else gamma, flag1, cpu_map, phi, f)
        self._chk(lib().rh_dump_snapshot(self._ctx, C.byref(t), C.byref(snapshot_info(**info)), os.fsencode(root)))

    def write_hydro(self, mesh: Mesh, path: str, gamma: float | None = None):
        """RamsesWriter::write_hydro of the bound mesh alone."""
        t, keep = mesh.snapshot_tree(self.params.gamma if gamma is None else gamma)
        self._chk(lib().rh_write_hydro(self._ctx, C.byref(t), C.byref(snapshot_info()), os.fsencode(path)))

    def courant(self, ilevel, dx, gamma, courant_factor) -> float:
        out = C.c_double()
        self._chk(lib().rh_courant(self._ctx, ilevel, dx, gamma, courant_factor, C.byref(out)))
        return out.value

    def courant_noalloc(self, ilevel, dx, gamma, courant_factor) -> float:
        out = C.c_double()
        self._chk(lib().rh_courant_noalloc(self._ctx, ilevel, dx, gamma, courant_factor, C.byref(out)))
        return out.value

    def run_uniform(self, ilevel, dx, courant_factor, nsteps) -> np.ndarray:
        dts = np.zeros(max(nsteps, 1), np.float64)
        self._chk(lib().rh_run_uniform(self._ctx, ilevel, dx, courant_factor, nsteps, _dp(dts)))
        return dts[:nsteps]

    def interpol(self, u1: np.ndarray, u2: np.ndarray | None = None) -> np.ndarray:
        u1 = np.ascontiguousarray(u1, np.float64)
        n = u1.shape[0]
        assert u1.shape == (n, 7, 64)
        u2 = np.zeros((n, 8, 64), np.float64) if u2 is None else np.ascontiguousarray(u2, np.float64)
        self._chk(lib().rh_interpol_hydro(self._ctx, n, _dp(u1), _dp(u2)))
        return u2

    def sync_to_host(self, what=RH_UOLD | RH_UNEW):
        self._chk(lib().rh_sync(self._ctx, RH_SYNC_TO_HOST, what))

    def sync_to_device(self, what=RH_UOLD | RH_UNEW):
        self._chk(lib().rh_sync(self._ctx, RH_SYNC_TO_DEVICE, what))

    def timers(self, enable=True):
        self._chk(lib().rh_timers_enable(self._ctx, 1 if enable else 0))

    def timer(self, kernel) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_longlong()
        self._chk(lib().rh_timers_read(self._ctx, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def timers_reset(self):
        self._chk(lib().rh_timers_reset(self._ctx))

    def autotune(self, ilevel, dt, dx, reps=3):
        best = C.c_int()
        ms = np.zeros(RH_KERNEL_VARIANTS, np.float64)
        self._chk(lib().rh_autotune(self._ctx, ilevel, dt, dx, reps, C.byref(best), _dp(ms)))
        return best.value, ms

    def variant(self) -> int:
        v = C.c_int()
        self._chk(lib().rh_get_kernel_variant(self._ctx, C.byref(v)))
        return v.value

    def set_unew_oct_lanes(self, ilevel):
        """set_unew with the Godunov kernel's lane mapping (counter
        calibration); the same result as set_unew."""
        self._chk(lib().rh_set_unew_oct_lanes(self._ctx, ilevel))

    def graphs(self, enable=True):
        """Replay multi-step exchanged-buffer runs from captured HIP graphs
        (the default) or launch every kernel eagerly."""
        self._chk(lib().rh_set_graphs(self._ctx, 1 if enable else 0))

    def graph_stats(self) -> tuple[int, int]:
        """(graphs captured, graph launches) of this context so far."""
        a, b = C.c_longlong(), C.c_longlong()
        self._chk(lib().rh_graph_stats(self._ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def leaf_cells(self, ilevel) -> int:
        n = C.c_longlong()
        self._chk(lib().rh_level_leaf_cells(self._ctx, ilevel, C.byref(n)))
        return n.value

    def close(self):
        if self._ctx:
            lib().rh_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AmrRun:
    """An adaptive run with the state on the GPU (rhydro_amr.h): built as
    Simulation::initialize builds it, then advanced by coarse steps."""

    def __init__(self, nx=1, levelmin=4, levelmax=6, ngridmax=20000, gamma=1.4, courant_factor=0.8,
                 err_grad_d=-1.0, err_grad_p=-1.0, err_grad_v=-1.0, nexpand=1, slope_type=1, riemann="hllc",
                 interpol_type=1, interpol_var=0, boxlen=1.0, regions=None, ic="regions", kh_amp=0.01,
                 precision=RH_PREC_FP64, device=0):
        regions = regions or []
        self._keep = []

        def col(key, default):
            a = np.array([float(r.get(key, default)) for r in regions] or [0.0], np.float64)
            self._keep.append(a)
            return _dp(a)

        types = np.array([1 if r.get("type", "cube") == "sphere" else 0 for r in regions] or [0], np.int32)
        self._keep.append(types)
        rv = riemann if isinstance(riemann, int) else resolve_riemann(riemann)
        self.params = RhParams(gamma, slope_type, rv, interpol_type, interpol_var, 5, 0, 0, boxlen, precision,
                               device)
        ap = RhAmrParams()
        ap.nx = ap.ny = ap.nz = nx
        ap.ngridmax, ap.levelmin, ap.levelmax = ngridmax, levelmin, levelmax
        ap.courant_factor = courant_factor
        ap.err_grad_d, ap.err_grad_p, ap.err_grad_v = err_grad_d, err_grad_p, err_grad_v
        ap.nexpand = nexpand
        ap.ic_kind = 1 if ic == "kh" else 0
        ap.nreg = len(regions)
        ap.reg_type = _ip(types)
        (ap.x_center, ap.y_center, ap.z_center, ap.length_x, ap.length_y, ap.length_z, ap.d_region, ap.u_region,
         ap.v_region, ap.w_region, ap.p_region) = (
            col("x_center", 0.5), col("y_center", 0.5), col("z_center", 0.5), col("length_x", 1.0),
            col("length_y", 1.0), col("length_z", 1.0), col("d", 1.0), col("u", 0.0), col("v", 0.0),
            col("w", 0.0), col("p", 1.0))
        ap.kh_amp = kh_amp
        self._h = C.c_void_p()
        st = lib().rh_amr_create(C.byref(self.params), C.byref(ap), C.byref(self._h))
        if st != RH_OK:
            raise RhError(st, lib().rh_last_error(None).decode())
        self.nlevelmax = levelmax

    def _chk(self, st):
        if st != RH_OK:
            raise RhError(st, lib().rh_last_error(lib().rh_amr_context(self._h)).decode())

    def step(self, n=1):
        self._chk(lib().rh_amr_step(self._h, n))

    def stats(self):
        t, n, cu = C.c_double(), C.c_int(), C.c_longlong()
        self._chk(lib().rh_amr_stats(self._h, C.byref(t), C.byref(n), C.byref(cu)))
        return t.value, n.value, cu.value

    def state(self) -> dict:
        """Copies of the run's arrays (state copied back from the GPU)."""
        m = RhMesh()
        nb, xg = _IP(), _DP()
        self._chk(lib().rh_amr_mesh(self._h, C.byref(m), C.byref(nb), C.byref(xg)))
        ng, nc, nl = m.ngridmax, m.ncell, m.nlevelmax

        def arr(p, n):
            return np.ctypeslib.as_array(p, shape=(n,)).copy()

        t, nstep, cu = self.stats()
        return dict(ncoarse=m.ncoarse, ngridmax=ng, ncell=nc, nlevelmax=nl, father=arr(m.father, ng),
                    son=arr(m.son, nc), nbor=arr(m.nbor, 6 * ng), next=arr(m.next, ng), headl=arr(m.headl, nl + 1),
                    numbl=arr(nb, nl + 1), xg=arr(xg, 3 * ng), uold=arr(m.uold, 5 * nc), unew=arr(m.unew, 5 * nc),
                    t=t, nstep=nstep, cell_updates=cu)

    def phase_ms(self) -> dict:
        ms = np.zeros(7, np.float64)
        self._chk(lib().rh_amr_timers(self._h, _dp(ms)))
        return dict(zip(("hydro", "flag", "make", "remove", "restrict", "tables", "flag_wait"), ms.tolist()))

    def solver_context(self):
        return lib().rh_amr_context(self._h)

    def dump(self, root: str, iout: int, **info):
        """dump_snapshot of the run into <root>/output_<iout>; `info` sets the
        SnapshotInfo fields the run does not supply (noutput, aexp, ...)."""
        self._chk(lib().rh_amr_dump(self._h, C.byref(snapshot_info(**info)), iout, os.fsencode(root)))

    def close(self):
        if self._h:
            lib().rh_amr_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
