"""Direct-hop overlay files in the reference's on-disk format (CPU): the name, and the zlib
stream of the int32 entries, against a file the reference seq build dumped."""

import os

import numpy as np

import oppx
from conftest import golden_path

REF_FILE = "dh_c384_gs1.2500E-01_ex0.0000E+00_0.0000E+00_0.0000E+00.bin"


def test_filename_matches_reference_format():
    assert oppx.dh_overlay_filename(384, 1 / 8) == REF_FILE
    assert oppx.dh_overlay_filename(48000, 1 / 32, (0.1, 0.2, 0.3)) == \
        "dh_c48000_gs3.1250E-02_ex1.0000E-01_2.0000E-01_3.0000E-01.bin"


def test_reference_written_file_loads_and_matches_builder(tmp_path):
    g = np.load(golden_path("fempic_small_dh_omp1.npz"), allow_pickle=False)
    n = g["out_overlay"].size
    ov = oppx.dh_overlay_load(golden_path(REF_FILE), n)
    assert np.array_equal(ov, g["out_overlay"])  # the reference's own dump of the same grid
    m = oppx.box_mesh(4, 4, 4)
    assert np.array_equal(ov, oppx.dh_overlay(m, (0, 0, 0), (1, 1, 1), 1 / 8))
    # our writer produces a file the same reader accepts, with the same entries
    path = str(tmp_path / REF_FILE)
    oppx.dh_overlay_save(path, ov)
    assert np.array_equal(oppx.dh_overlay_load(path, n), ov)
    with open(path, "rb") as a, open(golden_path(REF_FILE), "rb") as b:
        assert a.read() == b.read()  # zlib compress() of the same bytes
    # a grid of another size is refused
    try:
        oppx.dh_overlay_load(path, n + 1)
        raise AssertionError("accepted a size mismatch")
    except oppx.OppxError:
        pass
