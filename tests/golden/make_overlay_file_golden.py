"""Writes tests/golden/dh_c384_gs1.2500E-01_ex0.0000E+00_0.0000E+00_0.0000E+00.bin: the direct-hop
overlay file the REFERENCE seq build dumps (OPP_DH_DATA_DUMP=1) for the fempic_small_dh case
(4^3 cubes x 6 tets, spacing 1/8). Run where the reference tree exists:
    python tests/golden/make_overlay_file_golden.py
"""

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "op-pic_amd"))
sys.path.insert(0, HERE)
import make_fempic_golden as mg  # noqa: E402

NAME = "dh_c384_gs1.2500E-01_ex0.0000E+00_0.0000E+00_0.0000E+00.bin"


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "_ref/fempic_ref_seq"], check=True)
    m, ef, pos, vel, p2c = mg.build_inputs(4, 4000, 12, 0.02)
    prm = mg.case_params(4, 0.02)
    with tempfile.TemporaryDirectory() as td:
        ind, outd = os.path.join(td, "in"), os.path.join(td, "out")
        os.makedirs(ind)
        os.makedirs(outd)
        for k in ["node_pos", "c2n", "c2c", "c_det", "c_vol"]:
            m[k].tofile(os.path.join(ind, k + ".bin"))
        ef.tofile(os.path.join(ind, "c_ef.bin"))
        pos.tofile(os.path.join(ind, "p_pos.bin"))
        vel.tofile(os.path.join(ind, "p_vel.bin"))
        p2c.astype(np.int32).tofile(os.path.join(ind, "p2c.bin"))
        par = os.path.join(td, "case.param")
        mg.write_params(par, nn=m["n_nodes"], nc=m["n_cells"], npart=4000, steps=1, dt=prm["dt"],
                        charge=prm["charge"], mass=prm["mass"], dh=True, gs=1 / 8, bmin=(0, 0, 0), bmax=(1, 1, 1))
        subprocess.run([os.path.join(ROOT, "oracle", "_ref", "fempic_ref_seq"), par, ind, outd, "OPP_DH_DATA_DUMP=1"],
                       check=True, capture_output=True, cwd=td)
        shutil.copy(os.path.join(td, NAME), os.path.join(HERE, NAME))
    print("wrote", os.path.join(HERE, NAME))


if __name__ == "__main__":
    main()
