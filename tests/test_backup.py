"""Backup files in the reference's byte layout: exact bytes for a small mesh,
the restart reader's seek formula, and a round trip of an AMR Sedov state."""
import os
import struct

import numpy as np

from ramses2_amd.amr_build import sedov_amr
from ramses2_amd.backup import read_amr, read_backup, read_hydro, read_info, title, write_amr, write_backup, write_hydro


def test_title():
    assert title(1) == "00001" and title(12345) == "12345" and title(42) == "00042"


def test_amr_bytes(tmp_path):
    """ndim, levelmin, nlevelmax, noct per level, then (ckey, refined bits) per oct."""
    ck1 = np.array([[0, 0, 0]])
    ref1 = np.array([[1, 0, 0, 0, 0, 0, 0, 1]], dtype=bool)
    ck2 = np.array([[0, 0, 0], [1, 0, 0]])
    ref2 = np.zeros((2, 8), dtype=bool)
    p = tmp_path / "amr.00001"
    write_amr(str(p), 3, 1, 2, {1: (ck1, ref1), 2: (ck2, ref2)})
    want = struct.pack("<3i", 3, 1, 2) + struct.pack("<2i", 1, 2)
    want += struct.pack("<4i", 0, 0, 0, 1 | (1 << 7))
    want += struct.pack("<4i", 0, 0, 0, 0) + struct.pack("<4i", 1, 0, 0, 0)
    assert p.read_bytes() == want
    # the restart reader's seek: oct i at 1-based POS 13 + 4 nlev + (4 ndim + 4)(i-1)
    data = p.read_bytes()
    nlev, ndim = 2, 3
    for i, ck in enumerate([ck1[0], ck2[0], ck2[1]], start=1):
        pos = 13 + 4 * nlev + (4 * ndim + 4) * (i - 1)
        assert struct.unpack_from("<3i", data, pos - 1) == tuple(ck)


def test_hydro_bytes(tmp_path):
    u1 = np.arange(40, dtype=np.float64).reshape(1, 5, 8)
    p = tmp_path / "hydro.00001"
    write_hydro(str(p), 3, 4, 4, {4: u1})
    data = p.read_bytes()
    assert struct.unpack_from("<5i", data, 0) == (3, 5, 4, 4, 1)
    # uold(ind, ivar) with the cell index fastest: value ivar*8+ind at the
    # restart reader's POS 17 + 4 nlev + 8*(8*(ivar-1) + ind-1) (1-based)
    for ivar in range(1, 6):
        for ind in range(1, 9):
            pos = 17 + 4 * 1 + 8 * (8 * (ivar - 1) + ind - 1)
            assert struct.unpack_from("<d", data, pos - 1)[0] == u1[0, ivar - 1, ind - 1]
    assert read_hydro(str(p))[4][4].tolist() == u1.tolist()


def test_backup_round_trip(tmp_path):
    am, u = sedov_amr(levelmin=3, levelmax=5, radius=0.3)
    d = write_backup(str(tmp_path), 7, am, u, time=0.125, gamma=1.4, nstep_coarse=3)
    assert os.path.basename(d) == "backup_00007"
    info = read_info(os.path.join(d, "info.txt"))
    assert info["time"] == 0.125 and info["gamma"] == 1.4 and info["levelmin"] == 3 and info["levelmax"] == 5
    with open(os.path.join(d, "info.txt")) as f:
        lines = f.read().splitlines()
    assert lines[0] == "nfile       =          1"
    assert "time        =  0.125000000000000E+00" in lines
    am2, u2, _ = read_backup(str(tmp_path), 7)
    assert am2.levels == am.levels
    for l in am.levels:
        i0 = am.head[l] - 1
        k = am.ckey[i0:i0 + am.noct[l]]
        idx = am2.find(l, k)
        assert np.all(idx >= 0)
        assert np.array_equal(u2[idx], u[i0:i0 + am.noct[l]])
        assert np.array_equal(am2.refined[idx], am.refined[i0:i0 + am.noct[l]])
    ndim, lmin, lmax, lv = read_amr(os.path.join(d, "amr.00001"))
    assert (ndim, lmin, lmax) == (3, 3, 5) and sum(len(v[0]) for v in lv.values()) == am.ngrid
