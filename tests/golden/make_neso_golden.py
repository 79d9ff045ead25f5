"""Writes the NESO-advection golden fixtures from the REFERENCE build (oracle/_ref/neso_ref_seq).

Run where the reference tree exists:  python tests/golden/make_neso_golden.py
Inputs come from the oppx host builders; outputs only from the reference seq library and the
generated NESO-advection (mdir) seq loops. Fixtures are .npz (allow_pickle=False).
"""

from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "op-pic_amd"))
import oppx  # noqa: E402

GOLDEN = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(ROOT, "oracle", "_ref", "neso_ref_seq")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make_case(name, nx, n_part, seed, cw, dt, hops_per_step, steps, full_state=True):
    m = oppx.neso_mesh(nx, nx, cw)
    vmax = hops_per_step * cw / dt
    pos, vel, p2c = oppx.neso_particles(n_part, seed, m, vmax)
    with tempfile.TemporaryDirectory() as td:
        ind, outd = os.path.join(td, "in"), os.path.join(td, "out")
        os.makedirs(ind)
        os.makedirs(outd)
        m["c_pos_ll"].tofile(os.path.join(ind, "c_pos_ll.bin"))
        m["c2c"].tofile(os.path.join(ind, "c2c.bin"))
        pos.tofile(os.path.join(ind, "p_pos.bin"))
        vel.tofile(os.path.join(ind, "p_vel.bin"))
        p2c.tofile(os.path.join(ind, "p2c.bin"))
        par = os.path.join(td, "case.param")
        with open(par, "w") as f:
            f.write("\n".join([f"INT nx = {nx}", f"INT ny = {nx}", f"INT nparts = {n_part}", f"INT num_steps = {steps}",
                               f"REAL dt = {dt!r}", f"REAL cell_width = {cw!r}", "REAL opp_allocation_multiple = 1.1",
                               "BOOL opp_global_move = false", "STRING opp_fill = HoleFill_All",
                               "BOOL opp_profile_all = false"]) + "\n")
        r = subprocess.run([REF, par, ind, outd], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-2000:] + r.stdout[-2000:])
        rd = lambda n, t: np.fromfile(os.path.join(outd, n), dtype=t)
        out = dict(s1_pos=rd("s1_pos.bin", np.float64), s1_p2c=rd("s1_p2c.bin", np.int32),
                   s1_mdir=rd("s1_mdir.bin", np.int32), s1_hops=rd("s1_hops.bin", np.int32),
                   final_pos=rd("final_pos.bin", np.float64), final_p2c=rd("final_p2c.bin", np.int32),
                   final_mdir=rd("final_mdir.bin", np.int32))
        st = np.loadtxt(os.path.join(outd, "steps.txt"), ndmin=2).astype(np.int64)
    arrays = dict(meta_nx=np.array(nx), meta_n_part=np.array(n_part), meta_seed=np.array(seed), meta_cw=np.array(cw),
                  meta_dt=np.array(dt), meta_vmax=np.array(vmax), meta_steps=np.array(steps),
                  out_steps_iter_size_hops_maxhops=st)
    if full_state:
        arrays.update(in_pos=pos, in_vel=vel, in_p2c=p2c)
        arrays.update({f"out_{k}": v for k, v in out.items()})
    else:
        arrays.update({f"sha_{k}": np.array(sha(v)) for k, v in out.items()})
        arrays.update(sha_in_pos=np.array(sha(pos)), sha_in_vel=np.array(sha(vel)), sha_in_p2c=np.array(sha(p2c)))
    path = os.path.join(GOLDEN, f"neso_{name}.npz")
    np.savez_compressed(path, **arrays)
    print(f"{name}: {path} ({os.path.getsize(path)} B) steps {st[0].tolist()} .. {st[-1].tolist()}")


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "_ref/neso_ref_seq"], check=True)
    make_case("small", nx=16, n_part=5000, seed=3, cw=0.5, dt=0.5, hops_per_step=4.0, steps=10)
    make_case("far", nx=8, n_part=3000, seed=5, cw=0.5, dt=0.5, hops_per_step=11.0, steps=5)
    make_case("c3_mini", nx=256, n_part=1_000_000, seed=3, cw=0.5, dt=0.5, hops_per_step=4.0, steps=20,
              full_state=False)


if __name__ == "__main__":
    main()
