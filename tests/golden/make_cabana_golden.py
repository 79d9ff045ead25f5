"""Writes the CabanaPIC move-deposit fixtures from the REFERENCE build (oracle/_ref/cabana_ref_seq:
the reference seq backend and its generated seq move_deposit loop, on a periodic structured grid
with fixed interpolated fields).

Run in the container that holds the reference tree (it builds oracle/_ref from there):
    python tests/golden/make_cabana_golden.py
Inputs are seeded and generated here (numpy PCG64); outputs come only from the reference build.
Fixtures are numpy .npz files, loaded with allow_pickle=False.
"""

from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
GOLDEN = os.path.dirname(os.path.abspath(__file__))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "cabana_ref_seq")
sys.path.insert(0, GOLDEN)


def periodic_faces(nx, ny, nz):
    """Face neighbours {x-1, y-1, z-1, x+1, y+1, z+1} of cell x + nx*(y + ny*z), wrapped."""
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    x, y, z = x.T.ravel(), y.T.ravel(), z.T.ravel()  # cell order: x fastest
    cid = lambda a, b, c: (a % nx) + nx * ((b % ny) + ny * (c % nz))
    return np.stack([cid(x - 1, y, z), cid(x, y - 1, z), cid(x, y, z - 1), cid(x + 1, y, z), cid(x, y + 1, z),
                     cid(x, y, z + 1)], axis=1).astype(np.int32)


def cabana_inputs(nx, ny, nz, ppc, seed, vel_sigma):
    rng = np.random.default_rng(seed)
    nc = nx * ny * nz
    faces = periodic_faces(nx, ny, nz)
    interp = rng.normal(size=(nc, 18)) * np.array([0.5, 0.2, 0.2, 0.1] * 3 + [0.5, 0.2] * 3)
    n = nc * ppc
    p2c = np.repeat(np.arange(nc, dtype=np.int32), ppc)
    pos = rng.uniform(-1.0, 1.0, (n, 3))
    vel = rng.normal(0.0, vel_sigma, (n, 3))
    weight = rng.uniform(0.5, 1.5, n)
    return dict(n_cells=nc, c2ngc=faces, interp=interp, p2c=p2c, pos=pos, vel=vel, weight=weight,
                qsp=-1.0, qdt_2mc=-0.05, cdt_d=(0.7, 0.6, 0.5))


def run_ref(d, steps):
    with tempfile.TemporaryDirectory() as td:
        ind, outd = os.path.join(td, "in"), os.path.join(td, "out")
        os.makedirs(ind)
        os.makedirs(outd)
        for k, name in (("c2ngc", "c2ngc"), ("interp", "c_interp"), ("pos", "p_pos"), ("vel", "p_vel"),
                        ("weight", "p_weight"), ("p2c", "p2c")):
            np.ascontiguousarray(d[k]).tofile(os.path.join(ind, name + ".bin"))
        prm = [f"INT ncells = {d['n_cells']}", f"INT nparts = {d['p2c'].size}", f"INT num_steps = {steps}",
               f"REAL qsp = {d['qsp']!r}", f"REAL qdt_half_mc = {d['qdt_2mc']!r}", f"REAL cdt_dx = {d['cdt_d'][0]!r}",
               f"REAL cdt_dy = {d['cdt_d'][1]!r}", f"REAL cdt_dz = {d['cdt_d'][2]!r}",
               "REAL opp_allocation_multiple = 1.1", "BOOL opp_profile_all = false",
               "STRING opp_fill = HoleFill_All", "INT opp_fill_period = 4", "BOOL opp_global_move = false"]
        par = os.path.join(td, "case.param")
        with open(par, "w") as f:
            f.write("\n".join(prm) + "\n")
        r = subprocess.run([REF_BIN, par, ind, outd], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference cabana failed: {r.stderr[-2000:]}\n{r.stdout[-2000:]}")
        rd = lambda name, t: np.fromfile(os.path.join(outd, name + ".bin"), dtype=t)
        out = {}
        for t in ("s1", "final"):
            for k, dt_ in (("pos", np.float64), ("vel", np.float64), ("streak", np.float64), ("p2c", np.int32),
                           ("acc", np.float64), ("hops", np.int32)):
                out[f"{t}_{k}"] = rd(f"{t}_{k}", dt_)
        out["steps_iter_size_hops_maxhops"] = np.loadtxt(os.path.join(outd, "steps.txt"), ndmin=2).astype(np.int64)
        return out


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make_case(name, dims, ppc, seed, vel_sigma, steps, full_state):
    d = cabana_inputs(*dims, ppc, seed, vel_sigma)
    out = run_ref(d, steps)
    arrays = {f"meta_{k}": np.array(v) for k, v in dict(nx=dims[0], ny=dims[1], nz=dims[2], ppc=ppc, seed=seed,
                                                        vel_sigma=vel_sigma, steps=steps, qsp=d["qsp"],
                                                        qdt_2mc=d["qdt_2mc"]).items()}
    arrays["meta_cdt_d"] = np.array(d["cdt_d"])
    if full_state:
        arrays.update({f"in_{k}": d[k] for k in ("interp", "pos", "vel", "weight")})
        arrays.update({f"out_{k}": v for k, v in out.items()})
    else:
        for k in ("interp", "pos", "vel", "weight"):
            arrays[f"sha_in_{k}"] = np.array(sha(d[k]))
        for k, v in out.items():
            if k.endswith("acc") or k.startswith("steps"):
                arrays[f"out_{k}"] = v
            else:
                arrays[f"sha_{k}"] = np.array(sha(v))
    path = os.path.join(GOLDEN, f"cabana_{name}.npz")
    np.savez_compressed(path, **arrays)
    st = out["steps_iter_size_hops_maxhops"]
    print(f"{name}: wrote {path} ({os.path.getsize(path)} B); steps {st[0].tolist()} .. {st[-1].tolist()}")


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "_ref/cabana_ref_seq"], check=True)
    make_case("small", (6, 5, 4), 12, seed=51, vel_sigma=1.0, steps=4, full_state=True)
    make_case("grid16", (16, 16, 16), 40, seed=52, vel_sigma=0.8, steps=10, full_state=False)


if __name__ == "__main__":
    main()
