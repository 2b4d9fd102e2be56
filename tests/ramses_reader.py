"""Readers of RAMSES snapshot directories, for the snapshot tests.

`load_snapshot` restates how the reference's analysis reader
(tests/visu/visu_ramses.py:28-479, load_snapshot) finds its data in the amr_ and
hydro_ files: not by walking the records but by byte offsets
computed from counts of integers, reals, record frames and strings.  A file
that this reader loads correctly is one the reference's test suite loads.
Only what a one-rank, non-RT, particle-free snapshot needs is restated.

`records` walks a file of Fortran unformatted records and checks every frame.
"""
from __future__ import annotations

import math
import os
import struct

import numpy as np


def records(path: str) -> list[bytes]:
    """Payloads of the file's records; raises if a frame is inconsistent."""
    with open(path, "rb") as f:
        raw = f.read()
    out, pos = [], 0
    while pos < len(raw):
        (n,) = struct.unpack_from("<i", raw, pos)
        if n < 0 or pos + 8 + n > len(raw):
            raise ValueError(f"{path}: bad record length {n} at byte {pos}")
        (m,) = struct.unpack_from("<i", raw, pos + 4 + n)
        if m != n:
            raise ValueError(f"{path}: record at byte {pos} framed by {n} and {m}")
        out.append(raw[pos + 4: pos + 4 + n])
        pos += 8 + n
    return out


def read_info(path: str) -> dict:
    """key = value lines of info_NNNNN.txt (visu_ramses.py:38-47)."""
    info = {}
    with open(path) as f:
        for line in f:
            k, sep, v = line.partition("=")
            if not sep:
                continue
            v = v.strip()
            try:
                info[k.strip()] = int(v)
            except ValueError:
                try:
                    info[k.strip()] = float(v)
                except ValueError:
                    info[k.strip()] = v
    return info


def read_descriptor(path: str) -> list[str]:
    with open(path) as f:
        return [ln.split(",")[1].strip() for ln in f if not ln.startswith("#")]


def _off(ninteg, nfloat, nlines, nstrin=0):
    # Byte offset of a payload after ninteg ints, nfloat reals, nlines
    # record frames (8 bytes each) and nstrin string bytes, plus the leading
    # 4-byte count of the record itself.
    return 4 * ninteg + 8 * (nlines + nfloat) + nstrin + 4


def load_snapshot(root: str, iout: int) -> dict:
    """Cells of the snapshot as the reference's reader returns them: for every
    cell that is a leaf or sits on the last level, its hydro variables, level,
    x, y, z and dx (visu_ramses.py:73-78, 295-371).  The grav file is not
    read: the reference's reader expects the Fortran grav layout (4 header
    records, level/ncache records for every domain), which the reference's C++
    writer does not produce (5 header records plus the coarse phi and f, and
    level/ncache records only for domains with grids)."""
    num = str(iout).zfill(5)
    d = os.path.join(root, "output_" + num)
    info = read_info(os.path.join(d, f"info_{num}.txt"))
    names = read_descriptor(os.path.join(d, "hydro_file_descriptor.txt"))
    nvar = len(names)
    ndim, lmax, ncpu = info["ndim"], info["levelmax"], info["ncpu"]
    names = names + ["level", "x", "y", "z", "dx"]
    twotondim = 2 ** ndim

    def fname(kind):
        return os.path.join(d, f"{kind}_{num}.out{str(1).zfill(5)}")

    amr = open(fname("amr"), "rb").read()
    hyd = open(fname("hydro"), "rb").read()

    # Header fields at fixed offsets (visu_ramses.py:163-210).
    nx, ny, nz = struct.unpack_from("3i", amr, _off(2, 0, 2))
    nboundary = struct.unpack_from("i", amr, _off(7, 0, 5))[0]
    noutput = struct.unpack_from("i", amr, _off(9, 1, 8))[0]
    xbound = [float(nx // 2), float(ny // 2), float(nz // 2)]
    ngridlevel = np.zeros((ncpu + nboundary, lmax), np.int32)
    o = _off(14 + 2 * ncpu * lmax, 18 + 2 * noutput + 2 * lmax, 21)
    ngridlevel[:ncpu, :] = np.frombuffer(amr, np.int32, ncpu * lmax, o).reshape(lmax, ncpu).T
    if nboundary > 0:
        o = _off(14 + 3 * ncpu * lmax + 10 * lmax + 2 * nboundary * lmax, 18 + 2 * noutput + 2 * lmax, 25)
        ngridlevel[ncpu:, :] = np.frombuffer(amr, np.int32, nboundary * lmax, o).reshape(lmax, nboundary).T
    # Hilbert ordering: the bound-key record counts as a string whose length
    # is read from its own frame (visu_ramses.py:231-243).
    kpos = _off(14 + 3 * ncpu * lmax + 10 * lmax + 3 * nboundary * lmax + 5, 18 + 2 * noutput + 2 * lmax,
                21 + 2 + 3 * min(1, nboundary) + 1 + 1, 128) - 4
    key_size = struct.unpack_from("i", amr, kpos)[0]
    ncoarse = nx * ny * nz
    a_int = 14 + 3 * ncpu * lmax + 10 * lmax + 3 * nboundary * lmax + 5 + 3 * ncoarse
    a_flt = 18 + 2 * noutput + 2 * lmax
    a_lin = 21 + 2 + 3 * min(1, nboundary) + 1 + 1 + 1 + 3
    a_str = 128 + key_size
    h_int, h_flt, h_lin = 5, 1, 6

    pieces = []
    for il in range(lmax):
        dxcell = 0.5 ** (il + 1)
        xcent = np.zeros((twotondim, 3))
        for ind in range(twotondim):
            iz, iy, ix = ind // 4, (ind % 4) // 2, ind % 2
            xcent[ind] = ((ix - 0.5) * dxcell, (iy - 0.5) * dxcell, (iz - 0.5) * dxcell)
        for j in range(nboundary + ncpu):
            n = int(ngridlevel[j, il])
            h_lin += 2
            h_int += 2
            if n > 0:
                if j == 0:
                    xg = np.empty((n, ndim))
                    for k in range(ndim):
                        xg[:, k] = np.frombuffer(amr, np.float64, n,
                                                 _off(a_int + 3 * n, a_flt + k * (n + 1), a_lin + 3, a_str))
                    s_int, s_flt, s_lin = a_int + n * (4 + 2 * ndim), a_flt + n * ndim, a_lin + 4 + 3 * ndim
                    var = np.empty((n, twotondim, len(names)))
                    son = np.empty((n, twotondim), np.int32)
                    for ind in range(twotondim):
                        son[:, ind] = np.frombuffer(amr, np.int32, n,
                                                    _off(s_int + ind * n, s_flt, s_lin + ind, a_str))
                        for iv in range(nvar):
                            var[:, ind, iv] = np.frombuffer(
                                hyd, np.float64, n, _off(h_int, h_flt + (ind * nvar + iv) * (n + 1), h_lin))
                        var[:, ind, -5] = il + 1
                        for k in range(ndim):
                            var[:, ind, -4 + k] = (xg[:, k] + xcent[ind, k] - xbound[k]) * info["boxlen"]
                        var[:, ind, -1] = dxcell * info["boxlen"]
                    keep = ~((son > 0) & (il < lmax - 1))
                    pieces.append(var[keep])
                a_int += n * (4 + 3 * twotondim + 2 * ndim)
                a_flt += n * ndim
                a_lin += 4 + 3 * twotondim + 3 * ndim
                h_flt += n * twotondim * nvar
                h_lin += twotondim * nvar
    cells = np.concatenate(pieces) if pieces else np.zeros((0, len(names)))
    return {"info": info, "data": {k: cells[:, i] for i, k in enumerate(names)}, "ncells": len(cells)}


def solution_sums(data: dict) -> dict:
    """The aggregate of check_solution (visu_ramses.py:552-627): per variable,
    fsum of log10|v| for density and pressure and of |v| otherwise, after
    values within 1e-14 (relative) of the mean are set to the mean."""
    out = {}
    for key in sorted(data):
        v = np.asarray(data[key], np.float64)
        av = np.average(v)
        kv = v if av == 0.0 else np.where(np.abs(v - av) / abs(av) < 1.0e-14, av, v)
        sol = np.log10(np.abs(kv)) if key in ("density", "pressure") else np.abs(kv)
        out[key] = math.fsum(sol)
    return out
