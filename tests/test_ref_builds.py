"""The OpenMP builds of the reference's NESO-advection and CabanaPIC loops (oracle/_ref/*_ref_omp,
timed as CPU baselines) against its seq builds (the builds the golden fixtures come from), on the
same small inputs: final positions, cells, hop counts and move directions byte-identical, and the
CabanaPIC accumulators within 1e-14 relative (the omp kernel sums the same terms per cell in
another thread order). Skipped when the builds are absent (they are compiled from the reference
sources by oracle/Makefile)."""

import os
import subprocess

import numpy as np
import pytest

import oppx
from test_oracle_cabana import periodic_faces

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")


def run(binary, arrays, params, tmp, threads=None):
    """Runs oracle/_ref/<binary> on (array, stem, dtype) inputs and the `.param` lines; returns the
    output folder."""
    d = tmp / binary
    ind, outd = d / "in", d / "out"
    ind.mkdir(parents=True)
    outd.mkdir()
    for a, stem, dt in arrays:
        np.ascontiguousarray(a, dtype=dt).tofile(ind / f"{stem}.bin")
    (d / "case.param").write_text("\n".join(params) + "\n")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads)) if threads else None
    r = subprocess.run([os.path.join(REF, binary), str(d / "case.param"), str(ind), str(outd)], env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-400:]
    return outd


def need(*names):
    missing = [n for n in names if not os.path.exists(os.path.join(REF, n))]
    if missing:
        pytest.skip(f"reference builds absent: {missing}")


def same_bytes(a, b, name):
    return np.array_equal(np.fromfile(a / name, np.uint8), np.fromfile(b / name, np.uint8))


def test_neso_omp_build_matches_seq_build(tmp_path):
    need("neso_ref_seq", "neso_ref_omp")
    nx, cw, dt = 64, 0.5, 0.5
    m = oppx.neso_mesh(nx, nx, cw)
    pos, vel, p2c = oppx.neso_particles(100_000, 3, m, 4.0 * cw / dt)
    arrays = [(m["c_pos_ll"], "c_pos_ll", np.float64), (m["c2c"], "c2c", np.int32), (pos, "p_pos", np.float64),
              (vel, "p_vel", np.float64), (p2c, "p2c", np.int32)]
    params = [f"INT nx = {nx}", f"INT ny = {nx}", f"INT nparts = {len(p2c)}", "INT num_steps = 4",
              f"REAL dt = {dt!r}", f"REAL cell_width = {cw!r}", "REAL opp_allocation_multiple = 1.1",
              "BOOL opp_global_move = false", "STRING opp_fill = HoleFill_All", "BOOL opp_profile_all = false"]
    a = run("neso_ref_seq", arrays, params, tmp_path)
    b = run("neso_ref_omp", arrays, params, tmp_path, threads=4)
    for f in ("final_pos.bin", "final_p2c.bin", "final_mdir.bin", "s1_hops.bin"):
        assert same_bytes(a, b, f), f
    assert np.array_equal(np.loadtxt(a / "steps.txt", ndmin=2), np.loadtxt(b / "steps.txt", ndmin=2))


def test_cabana_omp_build_matches_seq_build(tmp_path):
    need("cabana_ref_seq", "cabana_ref_omp")
    nx, ny, nz, ppc = 8, 8, 10, 40
    nc = nx * ny * nz
    rng = np.random.default_rng(11)
    interp = rng.normal(size=(nc, 18)) * np.array([0.5, 0.2, 0.2, 0.1] * 3 + [0.5, 0.2] * 3)
    n = nc * ppc
    pos = rng.uniform(-1.0, 1.0, (n, 3))
    vel = rng.normal(0.0, 0.3, (n, 3))
    weight = rng.uniform(0.5, 1.5, n)
    p2c = np.repeat(np.arange(nc, dtype=np.int32), ppc)
    arrays = [(periodic_faces(nx, ny, nz), "c2ngc", np.int32), (interp, "c_interp", np.float64),
              (pos, "p_pos", np.float64), (vel, "p_vel", np.float64), (weight, "p_weight", np.float64),
              (p2c, "p2c", np.int32)]
    params = [f"INT ncells = {nc}", f"INT nparts = {n}", "INT num_steps = 3", "REAL qsp = -1.0",
              "REAL qdt_half_mc = -0.05", "REAL cdt_dx = 0.7", "REAL cdt_dy = 0.6", "REAL cdt_dz = 0.5",
              "REAL opp_allocation_multiple = 1.1", "BOOL opp_profile_all = false", "STRING opp_fill = HoleFill_All",
              "INT opp_fill_period = 4", "BOOL opp_global_move = false"]
    a = run("cabana_ref_seq", arrays, params, tmp_path)
    b = run("cabana_ref_omp", arrays, params, tmp_path, threads=4)
    for f in ("final_pos.bin", "final_vel.bin", "final_p2c.bin", "final_hops.bin"):
        assert same_bytes(a, b, f), f
    acc_a, acc_b = np.fromfile(a / "final_acc.bin"), np.fromfile(b / "final_acc.bin")
    assert np.abs(acc_a).max() > 0
    assert np.abs(acc_a - acc_b).max() <= 1e-14 * np.abs(acc_a).max()
