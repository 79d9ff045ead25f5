"""Digests of the reference's own HDF5 FEM-PIC meshes (scripts/batch/fempic/mesh_files_tar/box_*.tar.gz),
read with the pure-Python reader h5read.py: per mesh its box dimensions, and per dataset its dtype,
shape and the sha256 of its elements. Writes fempic_h5_meshes.json beside this script.

usage: python tests/golden/make_h5_mesh_digests.py [/root/reference]"""

import hashlib
import json
import os
import sys
import tarfile

import numpy as np

from h5read import H5

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
TARS = os.path.join(REF, "scripts/batch/fempic/mesh_files_tar")
# entries the reference's generator left uninitialised (heap contents at write time): dtype and shape only
UNINITIALISED = {"c_colors", "n_colors", "n_bnd_pot"}


def main():
    out = {}
    for name in ("box_24000", "box_48000", "box_96000"):
        with tarfile.open(os.path.join(TARS, name + ".tar.gz")) as t:
            data = t.extractfile(name + ".hdf5").read()
        f = H5(data)
        g = f.group(f.root)
        sets = {}
        for k, oh in sorted(g.items()):
            a = np.ascontiguousarray(f.dataset(oh))
            sets[k] = {"dtype": a.dtype.str, "shape": list(a.shape)}
            if k not in UNINITIALISED:
                sets[k]["sha256"] = hashlib.sha256(a.tobytes()).hexdigest()
        pos = f.dataset(g["n_pos"])
        # corner nodes sit at multiples of the cube edge, centre nodes half way between
        cube = 2.0 * float(np.diff(np.unique(pos[:, 0])).min())
        ext = pos.max(axis=0)
        dims = [int(round(float(e) / cube)) for e in ext]
        out[name] = {"file": f"scripts/batch/fempic/mesh_files_tar/{name}.tar.gz", "cube_mm": round(1000.0 * cube, 9),
                     "dims": dims,
                     "extent_m": [float(e) for e in ext], "datasets": sets}
        print(name, dims, cube, len(sets), "datasets")
    with open(os.path.join(HERE, "fempic_h5_meshes.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
