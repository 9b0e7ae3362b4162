"""Instant-field files of WRT_INSTANT_VARS_TG / RESTART_INSTANT_VARS_TG
(src/processing/WRT_VARS_TG.f90:1-91, src/processing/RESTART_tg.f90:1-46):
the reference's file names (Fortran ES15.9 edit of phyTIME), the 40-byte
header and the Fortran-order body, checked against byte strings assembled
independently with struct, and the restart's wall ghost rows."""
import os
import struct

import numpy as np
import pytest

from chapsim1ww_amd import fields
from chapsim1ww_amd.grid import ChannelGrid


@pytest.mark.parametrize("t,s", [(0.0, "0.000000000E+00"), (12.5, "1.250000000E+01"),
                                 (1234.56789, "1.234567890E+03"), (1e-3, "1.000000000E-03"),
                                 (0.123456789049, "1.234567890E-01"), (9.9999999996, "1.000000000E+01"),
                                 (7.0e42, "7.000000000E+42")])
def test_es15_9_edit(t, s):
    assert fields.es15_9(t) == s and len(s) == 15


def test_es15_9_refuses_what_does_not_fit():
    for t in (-1.0, 1e100, float("nan")):
        with pytest.raises(ValueError):
            fields.es15_9(t)


def test_reference_file_names():
    assert fields.instant_filename("U", 100.0) == "1_instant_D/DNS_perixz_INSTANT_T1.000000000E+02_U.D"
    assert fields.instant_filename("P", 0.5, "out/") == "out/DNS_perixz_INSTANT_T5.000000000E-01_P.D"
    with pytest.raises(ValueError):
        fields.instant_filename("T", 1.0)


def test_file_bytes_match_the_mpi_io_layout(tmp_path):
    """4 x int32 (NCL1, NCL2, NCL3, ITERG), 3 x float64 (phyTIME, REN, DT),
    then field(i, j, k) at i + NCL1 (j + NCL2 k), all little endian."""
    rng = np.random.default_rng(3)
    L, M, N = 5, 4, 3
    fld = np.asfortranarray(rng.standard_normal((L, M, N)))
    p = str(tmp_path / "x.D")
    fields.write_instant(p, fld, 1234, 2.5, 2800.0, 1e-3)
    expect = struct.pack("<4i3d", L, M, N, 1234, 2.5, 2800.0, 1e-3)
    expect += b"".join(struct.pack("<d", fld[i, j, k]) for k in range(N) for j in range(M) for i in range(L))
    assert open(p, "rb").read() == expect
    back, it, t, ren, dt = fields.read_instant(p)
    assert np.array_equal(back, fld) and (it, t, ren, dt) == (1234, 2.5, 2800.0, 1e-3)


def test_read_refuses_short_files(tmp_path):
    p = str(tmp_path / "y.D")
    open(p, "wb").write(struct.pack("<4i3d", 4, 4, 4, 1, 0.0, 1.0, 1.0) + b"\0" * 16)
    with pytest.raises(ValueError):
        fields.read_instant(p)
    open(p, "wb").write(b"\0" * 12)
    with pytest.raises(ValueError):
        fields.read_instant(p)


def test_write_and_restart_the_tg_state(tmp_path):
    g = ChannelGrid(16, 12, 8)
    q, pr, _ = fields.init_tg_fields(g, seed=5)
    rng = np.random.default_rng(9)
    pr[:, 1:g.M + 1, :] = rng.standard_normal((g.L, g.M, g.N))
    fields.apply_wall_bc(g, q, pr)
    d = str(tmp_path) + "/"
    paths = fields.write_instant_vars_tg(q, pr, 77, 3.25, g.ren, 2e-3, d)
    assert [os.path.basename(p) for p in paths] == [f"DNS_perixz_INSTANT_T3.250000000E+00_{c}.D" for c in "UVWP"]
    q2, pr2, it, t = fields.restart_instant_vars_tg(g, 3.25, d)
    assert (it, t) == (77, 3.25)
    assert np.array_equal(q2, q) and np.array_equal(pr2, pr)
    # the restart fills the ghost rows itself: u, w, v ghosts 0, v on the
    # bottom wall face 0, p ghosts the adjacent interior rows
    M = g.M
    assert np.all(q2[:, 0] == 0) and np.all(q2[:, M + 1] == 0) and np.all(q2[:, 1, :, 1] == 0)
    assert np.array_equal(pr2[:, 0], pr2[:, 1]) and np.array_equal(pr2[:, M + 1], pr2[:, M])
    with pytest.raises(ValueError):
        fields.restart_instant_vars_tg(ChannelGrid(16, 12, 16), 3.25, d)
