"""Writes the FEM-PIC golden fixtures from the REFERENCE build (oracle/_ref/fempic_ref_{seq,omp}).

Run in the container that holds the reference tree (it builds oracle/_ref from there):
    python tests/golden/make_fempic_golden.py
Inputs come from the oppx host builders (seeded, host independent); outputs come only from the
reference seq/omp libraries and generated FEM-PIC loops. Fixtures are numpy .npz files
(inputs + outputs), loaded with allow_pickle=False.
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
sys.path.insert(0, os.path.join(ROOT, "op-pic_amd"))
import oppx  # noqa: E402

GOLDEN = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def case_params(nx, sigma):
    h = 1.0 / nx
    dt = 0.5 * h / (sigma * math.sqrt(3.0))
    return dict(dt=dt, charge=1e-3, mass=1.0)


def build_inputs(nx, n_part, seed, sigma, sort_by_cell=False):
    m = oppx.box_mesh(nx, nx, nx)
    ef = oppx.box_field_sines(m)
    pos, vel, p2c = oppx.box_particles(n_part, seed, sigma, m)
    if sort_by_cell:
        pos, vel, p2c = oppx.sort_particles_by_cell(pos, vel, p2c)
    return m, ef, pos, vel, p2c


def write_params(path, *, nn, nc, npart, steps, dt, charge, mass, dh, gs, bmin, bmax, fill="HoleFill_All"):
    lines = [
        f"INT nnodes = {nn}", f"INT ncells = {nc}", f"INT nparts = {npart}", f"INT num_steps = {steps}",
        f"REAL dt = {dt!r}", f"REAL mass = {mass!r}", f"REAL charge = {charge!r}",
        f"REAL grid_spacing = {gs!r}",
        f"REAL bminx = {bmin[0]!r}", f"REAL bminy = {bmin[1]!r}", f"REAL bminz = {bmin[2]!r}",
        f"REAL bmaxx = {bmax[0]!r}", f"REAL bmaxy = {bmax[1]!r}", f"REAL bmaxz = {bmax[2]!r}",
        f"BOOL opp_global_move = {'true' if dh else 'false'}", "BOOL opp_dh_data_generate = true",
        "REAL opp_allocation_multiple = 1.1", "BOOL opp_profile_all = false",
        f"STRING opp_fill = {fill}", "INT opp_fill_period = 4",
    ]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def run_ref(binary, m, ef, pos, vel, p2c, *, steps, dt, charge, mass, dh, gs, bmin, bmax, threads=1):
    with tempfile.TemporaryDirectory() as td:
        ind, outd = os.path.join(td, "in"), os.path.join(td, "out")
        os.makedirs(ind)
        os.makedirs(outd)
        m["node_pos"].tofile(os.path.join(ind, "node_pos.bin"))
        m["c2n"].tofile(os.path.join(ind, "c2n.bin"))
        m["c2c"].tofile(os.path.join(ind, "c2c.bin"))
        m["c_det"].tofile(os.path.join(ind, "c_det.bin"))
        m["c_vol"].tofile(os.path.join(ind, "c_vol.bin"))
        ef.tofile(os.path.join(ind, "c_ef.bin"))
        pos.tofile(os.path.join(ind, "p_pos.bin"))
        vel.tofile(os.path.join(ind, "p_vel.bin"))
        p2c.astype(np.int32).tofile(os.path.join(ind, "p2c.bin"))
        par = os.path.join(td, "case.param")
        write_params(par, nn=m["n_nodes"], nc=m["n_cells"], npart=pos.shape[0], steps=steps, dt=dt, charge=charge,
                     mass=mass, dh=dh, gs=gs, bmin=bmin, bmax=bmax)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="false")
        r = subprocess.run([os.path.join(REF_DIR, binary), par, ind, outd], env=env, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{binary} failed: {r.stderr[-2000:]}\n{r.stdout[-2000:]}")
        out = {}
        rd = lambda name, dt_: np.fromfile(os.path.join(outd, name), dtype=dt_)
        for k, dt_ in [("s1_pos", np.float64), ("s1_vel", np.float64), ("s1_lc", np.float64), ("s1_p2c", np.int32),
                       ("s1_charge", np.float64), ("s1_hops", np.int32), ("final_pos", np.float64),
                       ("final_vel", np.float64), ("final_lc", np.float64), ("final_p2c", np.int32),
                       ("final_charge", np.float64), ("final_hops", np.int32)]:
            out[k] = rd(k + ".bin", dt_)
        if dh:
            out["overlay"] = rd("overlay.bin", np.int32)
            out["overlay_dims"] = rd("overlay_dims.bin", np.int64)
        st = np.loadtxt(os.path.join(outd, "steps.txt"), ndmin=2)
        out["steps_iter_before_after_hops_maxhops"] = st[:, :5].astype(np.int64)
        out["steps_qsum"] = st[:, 5].astype(np.float64)
        return out


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make_case(name, binary, nx, n_part, seed, sigma, steps, dh, threads, full_state=True, sort_by_cell=False,
              gs=None):
    m, ef, pos, vel, p2c = build_inputs(nx, n_part, seed, sigma, sort_by_cell)
    prm = case_params(nx, sigma)
    if gs is None:
        gs = 1.0 / (2 * nx) if dh else 0.25
    bmin, bmax = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    out = run_ref(binary, m, ef, pos, vel, p2c, steps=steps, gs=gs, bmin=bmin, bmax=bmax, dh=dh, threads=threads,
                  **prm)
    meta = dict(case=name, binary=binary, nx=nx, n_part=n_part, seed=seed, sigma=sigma, steps=steps, dh=int(dh),
                threads=threads, gs=gs, sort_by_cell=int(sort_by_cell), **prm)
    arrays = {f"meta_{k}": np.array(v) for k, v in meta.items()}
    arrays["bmin"] = np.array(bmin)
    arrays["bmax"] = np.array(bmax)
    if full_state:
        arrays.update({f"in_{k}": v for k, v in dict(pos=pos, vel=vel, p2c=p2c, c_ef=ef).items()})
        arrays.update({f"out_{k}": v for k, v in out.items()})
    else:
        # large case: digests of the per-particle arrays, full small arrays
        for k in ["final_pos", "final_vel", "final_lc", "final_p2c", "s1_p2c", "s1_hops", "final_hops"]:
            arrays[f"sha_{k}"] = np.array(sha(out[k]))
        for k in ["final_charge", "s1_charge", "steps_iter_before_after_hops_maxhops", "steps_qsum"]:
            arrays[f"out_{k}"] = out[k]
        if dh:
            arrays["out_overlay"] = out["overlay"]
            arrays["out_overlay_dims"] = out["overlay_dims"]
        arrays["sha_in_pos"] = np.array(sha(pos))
        arrays["sha_in_vel"] = np.array(sha(vel))
        arrays["sha_in_p2c"] = np.array(sha(p2c))
    path = os.path.join(GOLDEN, f"fempic_{name}.npz")
    np.savez_compressed(path, **arrays)
    st = out["steps_iter_before_after_hops_maxhops"]
    print(f"{name}: wrote {path} ({os.path.getsize(path)} B); steps {st[0].tolist()} .. {st[-1].tolist()}")


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "_ref/fempic_ref_seq", "_ref/fempic_ref_omp"],
                   check=True)
    make_case("tiny_seq", "fempic_ref_seq", nx=2, n_part=1000, seed=11, sigma=0.05, steps=10, dh=False, threads=1)
    make_case("small_dh_omp1", "fempic_ref_omp", nx=4, n_part=4000, seed=12, sigma=0.02, steps=10, dh=True, threads=1)
    make_case("small_dh_omp4", "fempic_ref_omp", nx=4, n_part=4000, seed=12, sigma=0.02, steps=10, dh=True, threads=4)
    make_case("c1", "fempic_ref_omp", nx=20, n_part=1_000_000, seed=1, sigma=0.01, steps=100, dh=True, threads=8,
              full_state=False, gs=1.0 / 32)


if __name__ == "__main__":
    main()
