"""Snapshots of GPU-resident runs (rhydro_snapshot.h) against the oracle's
restatement of the reference's writer: every file of output_NNNNN byte for
byte, for a uniform run, for a host-coherent binding whose host state was
edited just before the dump (density and pressure floors exercised), and for
an adaptive run; plus the reference's reader on the GPU's files."""
import filecmp
import os

import numpy as np
import pytest

import oracle
import ramses_reader
import rhydro

pytestmark = pytest.mark.gpu

FILES = ("amr_{n}.out00001", "hydro_{n}.out00001", "grav_{n}.out00001", "header_{n}.txt",
         "hydro_file_descriptor.txt", "part_file_descriptor.txt", "info_{n}.txt")


def _same_dirs(a, b, iout):
    n = str(iout).zfill(5)
    da, db = os.path.join(a, "output_" + n), os.path.join(b, "output_" + n)
    assert sorted(os.listdir(da)) == sorted(f.format(n=n) for f in FILES)
    for f in FILES:
        f = f.format(n=n)
        assert filecmp.cmp(os.path.join(da, f), os.path.join(db, f), shallow=False), f


def _oracle_dump(m, root, gamma, **info):
    arrays = {k: m.arrays[k] for k in ("headl", "taill", "numbl", "next", "prev", "father", "nbor", "son", "xg",
                                       "uold")}
    dims = dict(nx=m.nx, ny=m.ny, nz=m.nz, ngridmax=m.ngridmax, nlevelmax=m.nlevelmax, gamma=gamma,
                headf=m.headf, tailf=m.tailf, numbf=m.numbf)
    oracle.dump_snapshot(arrays, dims, root, **info)


def test_uniform_run_dump_equals_oracle(tmp_path):
    """Sedov 16^3 after 5 device-resident steps, with two empty levels above."""
    level = 5
    m = rhydro.uniform_mesh(level, nlevelmax=level + 2)
    rhydro.init_regions(m, rhydro.sedov_regions(level), 1.4)
    print("case snapshot uniform sedov16 5 steps", flush=True)
    s = rhydro.Solver(gamma=1.4, slope_type=1, riemann="hllc")
    s.bind(m)
    dts = s.run_uniform(level, 1.0 / 2 ** level, 0.8, 5)
    info = dict(t=float(dts.sum()), nstep=5, nstep_coarse=5, noutput=4, iout=3, levelmin=level,
                levelmax=level + 2, boxlen=1.0)
    g, o = tmp_path / "gpu", tmp_path / "oracle"
    g.mkdir()
    o.mkdir()
    s.dump_snapshot(m, str(g), **info)
    s.sync_to_host()
    _oracle_dump(m, str(o), 1.4, **info)
    _same_dirs(str(g), str(o), 3)
    snap = ramses_reader.load_snapshot(str(g), 3)
    assert snap["ncells"] == 8 ** (level - 1)
    leaf = m.cells_of_level(level, leaf_only=True)
    assert np.array_equal(snap["data"]["density"], np.maximum(m.state()[0, leaf], 1e-10))
    s.close()


def test_host_coherent_dump_floors(tmp_path):
    """Host edits reach the file (the dump uploads the host state first);
    rho below 1e-10 or negative is floored, a negative pressure clamped to 0,
    signed zeros kept."""
    level = 4
    m = rhydro.uniform_mesh(level)
    rhydro.init_lognormal(m, level, 11, 1.4)
    print("case snapshot host-coherent floors", flush=True)
    s = rhydro.Solver(gamma=1.4, slope_type=1, riemann="llf")
    s.bind(m, mode=rhydro.RH_HOST_COHERENT)
    s.godunov_fine(level, 1e-3, 1.0 / 2 ** level)
    u = m.state()
    leaf = m.cells_of_level(level, leaf_only=True)
    u[0, leaf[:4]] = [1e-12, -2.0, 0.0, 1e-10]
    u[4, leaf[4:8]] = [0.0, -1.0, 1e-300, -0.0]
    u[1, leaf[8]] = -0.0
    info = dict(t=0.5, nstep=1, nstep_coarse=1, noutput=1, iout=12, levelmin=level, levelmax=level)
    g, o = tmp_path / "gpu", tmp_path / "oracle"
    g.mkdir()
    o.mkdir()
    s.dump_snapshot(m, str(g), gamma=1.4, **info)
    _oracle_dump(m, str(o), 1.4, **info)
    _same_dirs(str(g), str(o), 12)
    snap = ramses_reader.load_snapshot(str(g), 12)
    assert snap["data"]["density"].min() == 1e-10 and snap["data"]["pressure"].min() == 0.0
    s.close()


@pytest.mark.parametrize("riemann", ["hllc"])
def test_amr_run_dump_equals_oracle(tmp_path, riemann):
    """Adaptive Sedov run (levels 4-6), 2 coarse steps, dumped by the GPU run
    and by the oracle's run: identical directories."""
    kw = dict(levelmin=4, levelmax=6, ngridmax=6000, err_grad_d=0.1, err_grad_p=0.1, regions=rhydro.sedov_regions(4))
    print(f"case snapshot amr sedov {riemann}", flush=True)
    g = rhydro.AmrRun(riemann=riemann, **kw)
    o = oracle.OracleAmr(riemann=rhydro.resolve_riemann(riemann), **kw)
    g.step(2)
    assert o.step(2) == 0
    dg, do = tmp_path / "gpu", tmp_path / "oracle"
    dg.mkdir()
    do.mkdir()
    g.dump(str(dg), 2, noutput=5, aexp=1.0)
    o.dump(str(do), 2, noutput=5, aexp=1.0)
    _same_dirs(str(dg), str(do), 2)
    snap = ramses_reader.load_snapshot(str(dg), 2)
    assert snap["ncells"] > 0 and np.isfinite(list(ramses_reader.solution_sums(snap["data"]).values())).all()
    g.close()


def test_write_hydro_alone_and_tree_mismatch(tmp_path):
    """rh_write_hydro writes the same bytes as the hydro file of a full dump;
    a tree that does not describe the bound mesh is rejected."""
    level = 4
    m = rhydro.uniform_mesh(level)
    rhydro.init_lognormal(m, level, 5, 1.4)
    print("case snapshot write_hydro alone", flush=True)
    s = rhydro.Solver(gamma=1.4)
    s.bind(m)
    s.write_hydro(m, str(tmp_path / "h.out"))
    s.dump_snapshot(m, str(tmp_path), iout=1)
    assert (tmp_path / "h.out").read_bytes() == (tmp_path / "output_00001" / "hydro_00001.out00001").read_bytes()
    other = rhydro.uniform_mesh(level + 1)
    with pytest.raises(rhydro.RhError):
        s.write_hydro(other, str(tmp_path / "bad.out"))
    s.close()
