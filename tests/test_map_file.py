"""The gridded integrated-parameter file (WRITE_INT_PAR_OUTPUT,
wam_output_module.f90:786-911) from the host entry point: its records
against an independent parser and against a Fortran program that reads it
with READ_GRID_FILE's statements (print/read_grid_file.f90, flang)."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import wam_amd

HERE = os.path.dirname(os.path.abspath(__file__))
CORNERS = [0, -16200000, 57564000, 16200000]     # AMOWEP, AMOSOP, AMOEAP, AMONOP [M_SEC]


def parse(path):
    recs = []
    with open(path, "rb") as fh:
        while True:
            head = fh.read(4)
            if not head:
                return recs
            n = struct.unpack("<i", head)[0]
            recs.append(fh.read(n))
            assert struct.unpack("<i", fh.read(4))[0] == n


def times(nx, ny, params, k):
    rng = np.random.default_rng(k)
    return [("2024010%d000000" % (i + 1), rng.random((len(params), ny, nx), dtype=np.float32)) for i in range(k)]


def test_map_records(tmp_path):
    nx, ny, params = 17, 9, [1, 2, 9, 11, 32]
    p = str(tmp_path / "MAP")
    sets = times(nx, ny, params, 2)
    for i, (d, g) in enumerate(sets):
        wam_amd.map_write(p, d, "20240101000000", nx, ny, CORNERS, [nx] * ny, [3600] * ny, params, g, append=i > 0)
    recs = parse(p)
    assert len(recs) == 2 * (3 + len(params))
    r = 0
    for d, g in sets:
        cd, fnx, fny, w, s, e, n, ca = struct.unpack("<14sff4i14s", recs[r])
        assert (cd.decode(), fnx, fny, [w, s, e, n], ca.decode()) == (d, nx, ny, CORNERS, "20240101000000")
        rows = np.frombuffer(recs[r + 1], np.int32)
        assert np.array_equal(rows, [nx] * ny + [3600] * ny)
        flags = np.frombuffer(recs[r + 2], np.int32)
        assert flags.size == 70 and list(np.flatnonzero(flags) + 1) == params and set(flags) == {0, 1}
        for j in range(len(params)):
            assert np.array_equal(np.frombuffer(recs[r + 3 + j], np.float32), g[j].ravel())
        r += 3 + len(params)
    with pytest.raises(ValueError):
        wam_amd.map_write(p, "x", "y", nx, ny, CORNERS, [nx] * ny, [3600] * ny, [2, 1], sets[0][1][:2])


def test_fortran_reads_map(tmp_path):
    fc = shutil.which("flang")
    if fc is None:
        pytest.skip("no Fortran compiler")
    exe = str(tmp_path / "map_io")
    subprocess.check_call([fc, "-O1", os.path.join(HERE, "fortran", "map_io.f90"), "-o", exe])
    nx, ny, params = 24, 13, [3, 5, 16, 32]
    p = str(tmp_path / "MAP")
    sets = times(nx, ny, params, 3)
    for i, (d, g) in enumerate(sets):
        wam_amd.map_write(p, d, "20231231120000", nx, ny, CORNERS, [nx] * ny, [7200] * ny, params, g, append=i > 0)
    subprocess.check_call([exe, p, str(len(sets))])
    raw = open(p + ".raw", "rb").read()
    o = 0
    for d, g in sets:
        cd, rnx, rny, w, s, e, n, ca = struct.unpack("<14s6i14s", raw[o:o + 52])
        assert (cd.decode(), rnx, rny, [w, s, e, n], ca.decode()) == (d, nx, ny, CORNERS, "20231231120000")
        o += 52
        rows = np.frombuffer(raw[o:o + 8 * ny], np.int32)
        assert np.array_equal(rows, [nx] * ny + [7200] * ny)
        o += 8 * ny
        flags = np.frombuffer(raw[o:o + 280], np.int32)
        assert list(np.flatnonzero(flags) + 1) == params
        o += 280
        for j in range(len(params)):
            assert np.array_equal(np.frombuffer(raw[o:o + 4 * nx * ny], np.float32), g[j].ravel())
            o += 4 * nx * ny
    assert o == len(raw)
