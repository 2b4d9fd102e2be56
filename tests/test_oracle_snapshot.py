"""Pins of the oracle's restatement of the reference's snapshot writer
(RamsesWriter.cpp, Simulation::dump_snapshot), and the library's host-side
amr file against it.

The pin is the reference's own reader: the files the oracle writes for an
adaptive run load through the byte-offset arithmetic of the reference's
visu_ramses.load_snapshot (restated in ramses_reader.py), every record is
framed consistently, and the cells that come back are the run's leaf cells,
level by level in list order, with exactly the primitive variables
(rho floored at 1e-10, u = rho u / rho, p = max((E - rho v^2/2)(gamma-1), 0))
the writer must produce.  A file with two records swapped fails the check."""
import os
import shutil

import numpy as np
import pytest

import oracle
import ramses_reader
import rhydro


def _lists(s):
    out = {}
    for L in range(1, s["nlevelmax"] + 1):
        g, seq = int(s["headl"][L]), []
        while g > 0:
            seq.append(g)
            g = int(s["next"][g - 1])
        out[L] = seq
    return out


def _expected_cells(s, gamma):
    """(level, rho, u, v, w, p) of the cells the reader returns, computed from
    the run's arrays without the writer: every oct of list L in list order,
    its 8 cells, kept when unrefined or on the last list."""
    nc, ng, lmax = s["ncoarse"], s["ngridmax"], s["nlevelmax"]
    u = s["uold"].reshape(5, -1)
    rows = []
    for L, octs in _lists(s).items():
        if not octs:
            continue
        g = np.array(octs, np.int64)
        cells = (nc + np.arange(8)[None, :] * ng + g[:, None] - 1).reshape(-1)  # oct-major, cell-minor
        keep = ~((s["son"][cells] > 0) & (L - 1 < lmax - 1))
        c = cells[keep]
        d = np.where(u[0, c] < 1e-10, 1e-10, u[0, c])
        vx, vy, vz = u[1, c] / d, u[2, c] / d, u[3, c] / d
        v2 = ((0.0 + vx * vx) + vy * vy) + vz * vz
        p = (u[4, c] - 0.5 * d * v2 - 0.0) * (gamma - 1.0)
        p = np.where(p < 0.0, 0.0, p)
        rows.append(np.stack([np.full(len(c), float(L)), d, vx, vy, vz, p], axis=1))
    return np.concatenate(rows)


def _check_loaded(root, iout, s, gamma):
    snap = ramses_reader.load_snapshot(root, iout)
    got = snap["data"]
    exp = _expected_cells(s, gamma)
    assert snap["ncells"] == len(exp)
    for k, name in enumerate(("level", "density", "velocity_x", "velocity_y", "velocity_z", "pressure")):
        assert np.array_equal(got[name].view(np.int64), exp[:, k].view(np.int64)), name
    return snap


@pytest.fixture(scope="module")
def amr_dump(tmp_path_factory):
    gamma = 1.4
    o = oracle.OracleAmr(levelmin=3, levelmax=5, ngridmax=3000, gamma=gamma, err_grad_d=0.1, err_grad_p=0.1,
                         riemann=2, regions=rhydro.sedov_regions(3))
    assert o.step(2) == 0
    root = str(tmp_path_factory.mktemp("snap"))
    o.dump(root, 7, noutput=3)
    return root, o.state(), gamma


def test_reader_loads_oracle_amr_dump(amr_dump):
    root, s, gamma = amr_dump
    assert s["numbl"][5] > 0 and s["numbl"][4] > 0, "the run must reach the last level"
    snap = _check_loaded(root, 7, s, gamma)
    info = snap["info"]
    assert info["ncpu"] == 1 and info["ndim"] == 3 and info["levelmax"] == 5 and info["levelmin"] == 3
    assert info["ngridmax"] == 3000 and info["nstep"] == s["nstep"] and info["noutput"] == 3
    assert info["time"] == s["t"] and info["boxlen"] == 1.0 and info["ordering type"] == "hilbert"


def test_every_record_is_framed(amr_dump):
    root, s, _ = amr_dump
    d = os.path.join(root, "output_00007")
    nonempty = sum(1 for L in range(1, s["nlevelmax"] + 1) if s["numbl"][L] > 0)
    amr = ramses_reader.records(os.path.join(d, "amr_00007.out00001"))
    hyd = ramses_reader.records(os.path.join(d, "hydro_00007.out00001"))
    grav = ramses_reader.records(os.path.join(d, "grav_00007.out00001"))
    # amr: 29 header records, then per non-empty level 37 (ids, next, prev,
    # 3 xg, father, 6 nbor, 8 son, 8 cpu_map, 8 flag1); hydro: 6, then per
    # level (il, ncache) and 40 value records when non-empty; grav: 5 + 4
    # coarse, then per non-empty level (il, ncache) + 32.
    assert len(amr) == 29 + 37 * nonempty
    assert len(hyd) == 6 + 2 * s["nlevelmax"] + 40 * nonempty
    assert len(grav) == 9 + 34 * nonempty
    # Records 20 (headl), 22 (numbl) and 25 (the ordering string).
    assert np.array_equal(np.frombuffer(amr[19], np.int32), s["headl"][1:])
    assert np.array_equal(np.frombuffer(amr[21], np.int32), s["numbl"][1:])
    assert amr[24].rstrip(b"\0") == b"hilbert" and len(amr[24]) == 128
    assert np.frombuffer(grav[5], np.float64).size == s["ncoarse"] and not np.frombuffer(grav[5]).any()


def test_swapped_records_fail_the_pin(amr_dump, tmp_path):
    """The pin is sensitive: swapping velocity_x and velocity_y of the first
    cell position in the last non-empty level breaks the loaded values."""
    root, s, gamma = amr_dump
    bad = tmp_path / "bad"
    shutil.copytree(root, bad)
    path = bad / "output_00007" / "hydro_00007.out00001"
    raw = bytearray(path.read_bytes())
    pos, spans = 0, []
    while pos < len(raw):
        n = int(np.frombuffer(raw, np.int32, 1, pos)[0])
        spans.append((pos, n))
        pos += 8 + n
    big = [sp for sp in spans if sp[1] > 8]
    (p1, n1), (p2, n2) = big[-40 + 1], big[-40 + 2]  # velocity_x, velocity_y of cell 1
    assert n1 == n2 and not np.array_equal(raw[p1:p1 + n1], raw[p2:p2 + n2])
    a, b = bytes(raw[p1 + 4:p1 + 4 + n1]), bytes(raw[p2 + 4:p2 + 4 + n2])
    raw[p1 + 4:p1 + 4 + n1], raw[p2 + 4:p2 + 4 + n2] = b, a
    path.write_bytes(bytes(raw))
    with pytest.raises(AssertionError):
        _check_loaded(str(bad), 7, s, gamma)


def test_solution_sums_of_oracle_dump(amr_dump):
    """check_solution's aggregate over the loaded cells is finite and counts
    every variable the reference's tests compare."""
    root, s, _ = amr_dump
    sums = ramses_reader.solution_sums(ramses_reader.load_snapshot(root, 7)["data"])
    assert set(sums) == {"density", "velocity_x", "velocity_y", "velocity_z", "pressure", "level", "x", "y", "z",
                         "dx"}
    assert all(np.isfinite(v) for v in sums.values())


@pytest.mark.parametrize("levelmin,nlevelmax", [(3, 3), (4, 6)])
def test_library_amr_file_equals_oracle(tmp_path, levelmin, nlevelmax):
    """The library writes the amr file on the host (rh_write_amr): byte for byte
    the oracle's, for uniform meshes with and without empty levels."""
    m = rhydro.uniform_mesh(levelmin, nlevelmax=nlevelmax, ngridmax=int(rhydro.lib().rh_uniform_ngrid(1, 1, 1,
                                                                                                        levelmin)) + 5)
    rng = np.random.default_rng(3)
    flag1 = rng.integers(0, 2, m.ncell).astype(np.int32)
    info = dict(t=0.125, nstep=4, nstep_coarse=4, noutput=2, iout=5, levelmin=levelmin, levelmax=nlevelmax,
                boxlen=2.0, aexp=0.5, omega_m=0.3)
    lib_path = tmp_path / "amr_lib.out"
    rhydro.write_amr(m, str(lib_path), flag1=flag1, **info)
    arrays = {k: m.arrays[k] for k in ("headl", "taill", "numbl", "next", "prev", "father", "nbor", "son", "xg",
                                       "uold")}
    arrays["flag1"] = flag1
    dims = dict(nx=1, ny=1, nz=1, ngridmax=m.ngridmax, nlevelmax=nlevelmax, gamma=1.4, headf=m.headf,
                tailf=m.tailf, numbf=m.numbf)
    oracle.dump_snapshot(arrays, dims, str(tmp_path), **info)
    ref = (tmp_path / "output_00005" / "amr_00005.out00001").read_bytes()
    assert lib_path.read_bytes() == ref
    mem = np.frombuffer(ramses_reader.records(str(lib_path))[23], np.int32)
    assert list(mem) == [m.headf, m.tailf, m.numbf, 0, 0]
