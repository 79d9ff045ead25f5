"""Writes the fixtures of the 12-tet box meshes from the REFERENCE builds.

Run where the reference tree exists:  python tests/golden/make_mesh12_golden.py

* mesh12_loader.npz: what the reference app's mesh loaders (oracle/_ref/mesh_ref: LoadVolumeMesh and
  LoadSurfaceMesh of apps/app_fempic/minifempic_funcs.h) compute from mesh files: in full for a small
  mesh whose node coordinates are jittered (the files are kept beside it, tests/golden/mesh12_jitter/),
  as digests for the 20 x 20 x 10-cube, 48,000-tet box (the size of the reference's box_48000 mesh).
  The files are written by oppx_box12_write (the jitter by this script, seed 3); only the arrays come
  from the reference build.
* fempic_mesh12_dh.npz: the reference FEM-PIC hot path (oracle/_ref/fempic_ref_omp on one thread: push,
  direct-hop move, deposit, hole fill) on a 6 x 6 x 5-cube 12-tet box read through the loader, 8000 particles,
  10 steps, with the reference's own direct-hop overlay. Inputs come from the oppx host builders.
"""

from __future__ import annotations

import hashlib
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
GOLDEN = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "op-pic_amd"))
sys.path.insert(0, GOLDEN)
import oppx  # noqa: E402
from make_fempic_golden import run_ref  # noqa: E402

MESH_REF = os.path.join(ROOT, "oracle", "_ref", "mesh_ref")
JITTER_DIR = os.path.join(GOLDEN, "mesh12_jitter")
LOADER_KEYS = [("node_pos", np.float64), ("node_vol", np.float64), ("node_type", np.int32), ("c2n", np.int32),
               ("c2c", np.int32), ("c_det", np.float64), ("c_vol", np.float64), ("if2c", np.int32),
               ("if2n", np.int32), ("if_u", np.float64), ("if_v", np.float64), ("if_normal", np.float64),
               ("if_area", np.float64)]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def reference_loader(d):
    with tempfile.TemporaryDirectory() as td:
        subprocess.run([MESH_REF, f"{d}/mesh.dat", f"{d}/inlet.dat", f"{d}/wall.dat", td], check=True,
                       capture_output=True)
        return {k: np.fromfile(os.path.join(td, k + ".bin"), dtype=t) for k, t in LOADER_KEYS}


def write_jittered(d, nx, ny, nz, cube_mm, jitter_mm, seed):
    oppx.box12_write(nx, ny, nz, cube_mm, d)
    lines = open(os.path.join(d, "mesh.dat")).read().splitlines()
    nn = int(lines[0].split()[0])
    rng = np.random.default_rng(seed)
    out = [lines[0]]
    for ln in lines[1:1 + nn]:
        i, x, y, z = ln.split()
        j = rng.uniform(-jitter_mm, jitter_mm, 3)
        out.append(f"{i} {float(x) + j[0]:.14e} {float(y) + j[1]:.14e} {float(z) + j[2]:.14e}")
    with open(os.path.join(d, "mesh.dat"), "w") as f:
        f.write("\n".join(out + lines[1 + nn:]) + "\n")


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "_ref/mesh_ref", "_ref/fempic_ref_omp"], check=True)
    # ---- loader fixtures
    os.makedirs(JITTER_DIR, exist_ok=True)
    write_jittered(JITTER_DIR, 6, 5, 4, 33.3, 3.3, 3)
    arrays = {f"jitter_{k}": v for k, v in reference_loader(JITTER_DIR).items()}
    with tempfile.TemporaryDirectory() as td:
        oppx.box12_write(20, 20, 10, 50.0, td)
        big = reference_loader(td)
    arrays.update({f"box48k_sha_{k}": np.array(sha(v)) for k, v in big.items()})
    arrays["box48k_sizes"] = np.array([len(big["node_vol"]), len(big["c_vol"]), len(big["if_area"])])
    np.savez_compressed(os.path.join(GOLDEN, "mesh12_loader.npz"), **arrays)
    print("mesh12_loader: jitter", len(arrays["jitter_c_vol"]), "tets; box48k", arrays["box48k_sizes"].tolist())

    # ---- hot path on a 12-tet box read through the loader
    nx, ny, nz, n_part, seed, sigma, steps = 6, 6, 5, 8000, 31, 0.02, 10
    m = oppx.box12_mesh(nx, ny, nz, 1.0, 1.0, nz / nx)
    ef = oppx.box_field_sines(m)
    pos, vel, p2c = oppx.box12_particles(n_part, seed, sigma, m)
    pos, vel, p2c = oppx.sort_particles_by_cell(pos, vel, p2c)
    dt = 0.5 * (1.0 / nx) / (sigma * math.sqrt(3.0))
    gs = 1.0 / (2 * nx)
    bmin, bmax = (0.0, 0.0, 0.0), (1.0, 1.0, nz / nx)
    # the omp backend on one thread: its direct-hop move removes particles outside the overlay cuboid
    # (opp_lib/include/opp_omp.h:203-292), where the seq backend reads out of bounds (opp_seq.h:91-124)
    out = run_ref("fempic_ref_omp", m, ef, pos, vel, p2c, steps=steps, dt=dt, charge=1e-3, mass=1.0, dh=True, gs=gs,
                  bmin=bmin, bmax=bmax, threads=1)
    meta = dict(case="mesh12_dh", binary="fempic_ref_omp", nx=nx, ny=ny, nz=nz, n_part=n_part, seed=seed, sigma=sigma,
                steps=steps, dh=1, gs=gs, dt=dt, charge=1e-3, mass=1.0)
    fx = {f"meta_{k}": np.array(v) for k, v in meta.items()}
    fx.update(bmin=np.array(bmin), bmax=np.array(bmax))
    fx.update({f"in_{k}": v for k, v in dict(pos=pos, vel=vel, p2c=p2c, c_ef=ef).items()})
    fx.update({f"out_{k}": v for k, v in out.items()})
    np.savez_compressed(os.path.join(GOLDEN, "fempic_mesh12_dh.npz"), **fx)
    st = out["steps_iter_before_after_hops_maxhops"]
    print(f"fempic_mesh12_dh: {m['n_cells']} tets; steps {st[0].tolist()} .. {st[-1].tolist()}")


if __name__ == "__main__":
    main()
