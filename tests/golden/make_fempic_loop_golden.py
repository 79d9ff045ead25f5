"""Writes the FEM-PIC main-loop fixtures (injection and the loops either side of the field solve)
from the REFERENCE build (oracle/_ref/fempic_loop_ref_seq: the reference seq backend and its
generated seq loops, apps/app_fempic/fempic.cpp:138-231 with the PETSc solve replaced by a fixed
node potential).

Run in the container that holds the reference tree (it builds oracle/_ref from there):
    python tests/golden/make_fempic_loop_golden.py
Inputs come from the oppx host builders plus the synthetic inlet below (seeded, host
independent); outputs come only from the reference build. Fixtures are numpy .npz files, loaded
with allow_pickle=False.
"""

from __future__ import annotations

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
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "fempic_loop_ref_seq")
INJ_EXCESS = 100  # random pairs beyond one step's injection (apps/app_fempic/fempic_defs.h:50)


def loop_inputs(nx: int, seed: int, per_face_max: int, sigma: float = 0.02):
    """Mesh, field-side dats and an inlet made of the boundary faces on z = 0."""
    m = oppx.box_mesh(nx, nx, nx)
    ef = oppx.box_field_sines(m)
    rng = np.random.default_rng(seed)
    npos, c2n, c2c = m["node_pos"].reshape(-1, 3), m["c2n"].reshape(-1, 4), m["c2c"].reshape(-1, 4)
    det, vol = m["c_det"].reshape(-1, 4, 4), m["c_vol"]
    # shape derivatives: grad lc_k = (-b_k, g_k, -d_k) / (6V), 12 doubles per tet
    s = 1.0 / (6.0 * vol)
    sd = np.stack([-det[:, :, 1] * s[:, None], det[:, :, 2] * s[:, None], -det[:, :, 3] * s[:, None]], axis=2)
    sd = np.ascontiguousarray(sd.reshape(-1, 12))
    n_vol = np.zeros(m["n_nodes"])
    np.add.at(n_vol, c2n.ravel(), np.repeat(vol / 4.0, 4))
    z = npos[:, 2]
    n_type = np.where(z == 1.0, 3, np.where(z == 0.0, 2, 0)).astype(np.int32)
    n_pot = np.sin(math.pi * npos[:, 0]) * np.sin(math.pi * npos[:, 1]) * np.sin(math.pi * npos[:, 2])
    n_pot = n_pot + 0.25 * npos[:, 2]
    faces = []
    for c in range(m["n_cells"]):
        for v in range(4):
            if c2c[c, v] != -1:
                continue
            nodes = [c2n[c, w] for w in range(4) if w != v]
            if all(z[n] == 0.0 for n in nodes):
                faces.append((c, nodes))
    if2c = np.array([c for c, _ in faces], np.int32)
    p = np.array([[npos[n] for n in nodes] for _, nodes in faces])  # [F][3][3]
    u, w = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]
    normal = np.tile(np.array([0.0, 0.0, 1.0]), (len(faces), 1))
    counts = rng.integers(0, per_face_max + 1, len(faces))
    counts[1] = 0  # a face without particles: the distribution's cumulative count repeats
    dist = np.cumsum(counts).astype(np.int32)
    inject = int(dist[-1])
    rand = rng.random((inject + INJ_EXCESS, 2))
    h = 1.0 / nx
    dt = 0.5 * h / (sigma * math.sqrt(3.0))
    return dict(mesh=m, ef=ef, sd=sd, n_vol=n_vol, n_type=n_type, n_pot=n_pot, if2c=if2c,
                if_u=np.ascontiguousarray(u), if_v=np.ascontiguousarray(w), if_norm=normal,
                if_npos=np.ascontiguousarray(p.reshape(-1, 9)), if_dist=dist, rand=rand, inject=inject,
                dt=dt, ion_velocity=0.4 * h / dt, charge=1e-3, mass=1.0, spwt=2.0, wall_potential=1.5)


def loop_inputs_mesh12(nx: int, ny: int, nz: int, seed: int, per_face_max: int, sigma: float = 0.02):
    """The 12-tet box of nx x ny x nz cubes written in the app's text format and read back through the
    loader (oppx.box12_mesh: the geometry, node volumes, node types and inlet faces the reference app's
    LoadVolumeMesh / LoadSurfaceMesh compute from those files); the field-side dats and the inlet
    distribution synthetic as in loop_inputs."""
    m = oppx.box12_mesh(nx, ny, nz, 1.0, ny / nx, nz / nx)
    ef = oppx.box_field_sines(m)
    rng = np.random.default_rng(seed)
    npos, c2n = m["node_pos"].reshape(-1, 3), m["c2n"].reshape(-1, 4)
    det, vol = m["c_det"].reshape(-1, 4, 4), m["c_vol"]
    s = 1.0 / (6.0 * vol)
    sd = np.stack([-det[:, :, 1] * s[:, None], det[:, :, 2] * s[:, None], -det[:, :, 3] * s[:, None]], axis=2)
    sd = np.ascontiguousarray(sd.reshape(-1, 12))
    n_pot = np.sin(math.pi * npos[:, 0]) * np.sin(math.pi * npos[:, 1]) * np.sin(math.pi * npos[:, 2])
    n_pot = n_pot + 0.25 * npos[:, 2]
    nf = len(m["if2c"])
    counts = rng.integers(0, per_face_max + 1, nf)
    counts[1] = 0  # a face without particles: the distribution's cumulative count repeats
    dist = np.cumsum(counts).astype(np.int32)
    inject = int(dist[-1])
    rand = rng.random((inject + INJ_EXCESS, 2))
    h = 1.0 / nx
    dt = 0.5 * h / (sigma * math.sqrt(3.0))
    if_npos = np.ascontiguousarray(npos[m["if2n"].reshape(-1, 3)].reshape(-1, 9))
    return dict(mesh=m, ef=ef, sd=sd, n_vol=m["node_vol"], n_type=m["node_type"], n_pot=n_pot, if2c=m["if2c"],
                if_u=np.ascontiguousarray(m["if_u"]), if_v=np.ascontiguousarray(m["if_v"]),
                if_norm=np.ascontiguousarray(m["if_normal"]), if_npos=if_npos, if_dist=dist, rand=rand,
                inject=inject, dt=dt, ion_velocity=0.4 * h / dt, charge=1e-3, mass=1.0, spwt=2.0, wall_potential=1.5,
                bmax=(1.0, ny / nx, nz / nx), mesh12=(nx, ny, nz))


def run_ref(d, *, steps, dh, gs):
    m = d["mesh"]
    with tempfile.TemporaryDirectory() as td:
        ind, outd = os.path.join(td, "in"), os.path.join(td, "out")
        os.makedirs(ind)
        os.makedirs(outd)
        files = dict(node_pos=m["node_pos"], c2n=m["c2n"], c2c=m["c2c"], c_det=m["c_det"], c_vol=m["c_vol"],
                     c_ef=d["ef"], c_sd=d["sd"], n_vol=d["n_vol"], n_type=d["n_type"], n_pot=d["n_pot"],
                     if2c=d["if2c"], if_u=d["if_u"], if_v=d["if_v"], if_norm=d["if_norm"], if_npos=d["if_npos"],
                     if_dist=d["if_dist"], rand=d["rand"])
        for k, a in files.items():
            np.ascontiguousarray(a).tofile(os.path.join(ind, k + ".bin"))
        prm = [f"INT nnodes = {m['n_nodes']}", f"INT ncells = {m['n_cells']}", f"INT nfaces = {len(d['if2c'])}",
               f"INT nrand = {d['rand'].shape[0]}", f"INT inject_count = {d['inject']}", f"INT num_steps = {steps}",
               f"REAL dt = {d['dt']!r}", f"REAL mass = {d['mass']!r}", f"REAL charge = {d['charge']!r}",
               f"REAL spwt = {d['spwt']!r}", f"REAL ion_velocity = {d['ion_velocity']!r}",
               f"REAL wall_potential = {d['wall_potential']!r}", f"REAL grid_spacing = {gs!r}",
               "REAL bminx = 0.0", "REAL bminy = 0.0", "REAL bminz = 0.0",
               *[f"REAL bmax{a} = {v!r}" for a, v in zip("xyz", d.get("bmax", (1.0, 1.0, 1.0)))],
               f"BOOL opp_global_move = {'true' if dh else 'false'}", "BOOL opp_dh_data_generate = true",
               "REAL opp_allocation_multiple = 1.1", "BOOL opp_profile_all = false",
               "STRING opp_fill = HoleFill_All", "INT opp_fill_period = 4"]
        par = os.path.join(td, "case.param")
        with open(par, "w") as f:
            f.write("\n".join(prm) + "\n")
        r = subprocess.run([REF_BIN, par, ind, outd], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference loop failed: {r.stderr[-2000:]}\n{r.stdout[-2000:]}")
        rd = lambda name, t: np.fromfile(os.path.join(outd, name + ".bin"), dtype=t)
        out = {"bnd_pot": rd("bnd_pot", np.float64), "dist1_p2c": rd("dist1_p2c", np.int32),
               "inj1_pos": rd("inj1_pos", np.float64), "inj1_vel": rd("inj1_vel", np.float64),
               "inj1_p2c": rd("inj1_p2c", np.int32)}
        for t in ("s1", "final"):
            for k, dt_ in (("pos", np.float64), ("vel", np.float64), ("lc", np.float64), ("p2c", np.int32),
                           ("charge", np.float64), ("ef", np.float64)):
                out[f"{t}_{k}"] = rd(f"{t}_{k}", dt_)
        rows = [ln.split() for ln in open(os.path.join(outd, "steps.txt")).read().splitlines() if ln.strip()]
        out["steps_iter_before_after_hops_maxhops"] = np.array([[int(x) for x in r[:5]] for r in rows], np.int64)
        out["steps_max_cef_chg_pot"] = np.array([[float.fromhex(x) for x in r[5:8]] for r in rows], np.float64)
        return out


def make_case(name, nx, seed, per_face_max, steps, dh, mesh12=None):
    d = loop_inputs(nx, seed, per_face_max) if mesh12 is None else loop_inputs_mesh12(*mesh12, seed, per_face_max)
    gs = 1.0 / (2 * nx)
    out = run_ref(d, steps=steps, dh=dh, gs=gs)
    arrays = {f"meta_{k}": np.array(v) for k, v in dict(nx=nx, seed=seed, per_face_max=per_face_max, steps=steps,
                                                        dh=int(dh), gs=gs, dt=d["dt"], ion_velocity=d["ion_velocity"],
                                                        charge=d["charge"], mass=d["mass"], spwt=d["spwt"],
                                                        wall_potential=d["wall_potential"],
                                                        inject=d["inject"]).items()}
    if mesh12 is not None:  # the 12-tet box read through the mesh-file loader
        arrays.update(meta_mesh12=np.array(mesh12), meta_bmax=np.array(d["bmax"]))
    for k in ("sd", "n_vol", "n_type", "n_pot", "if2c", "if_u", "if_v", "if_norm", "if_npos", "if_dist", "rand"):
        arrays[f"in_{k}"] = np.ascontiguousarray(d[k])
    arrays.update({f"out_{k}": v for k, v in out.items()})
    path = os.path.join(GOLDEN, f"fempic_loop_{name}.npz")
    np.savez_compressed(path, **arrays)
    st = out["steps_iter_before_after_hops_maxhops"]
    print(f"{name}: wrote {path} ({os.path.getsize(path)} B); faces {len(d['if2c'])}, inject {d['inject']}/step; "
          f"steps {st[0].tolist()} .. {st[-1].tolist()}")


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "_ref/fempic_loop_ref_seq"], check=True)
    make_case("small", nx=4, seed=31, per_face_max=12, steps=8, dh=False)
    make_case("dh", nx=6, seed=32, per_face_max=20, steps=12, dh=True)
    make_case("mesh12", nx=6, seed=33, per_face_max=20, steps=12, dh=True, mesh12=(6, 6, 5))


if __name__ == "__main__":
    main()
