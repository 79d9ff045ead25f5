"""GPU parity of the opp_lib drop-in (libopp_mi355x.so): applications written only against the
opp_lib API (tests/api/*.cpp, shaped like OP-PIC's generated mains
app_fempic_cg/fempic_opp.cpp:52-231 and app_neso_advection_mdir_cg/advec_opp.cpp:50-129) run on the
GPU and must reproduce the reference fixtures: cell indices, sizes, max hops, positions, lc and the
direct-hop overlay bit-exact, node charge within 1e-12 relative (FP64 atomic reordering)."""

import os
import subprocess
import tempfile

import numpy as np
import pytest

import oppx
from conftest import golden_path

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
API_DIR = os.path.join(ROOT, "tests", "api")
DEPOSIT_RTOL = 1e-12


def app(name):
    path = os.path.join(API_DIR, name)
    if not os.path.exists(path):  # host-only link step (g++), no device code is compiled here
        subprocess.run(["make", "-C", API_DIR, name], check=True, capture_output=True)
    return path


def run_app(binary, params, ind, outd):
    print(f"[api] {os.path.basename(binary)} {params}", flush=True)
    r = subprocess.run([binary, params, ind, outd], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, f"{binary} rc={r.returncode}\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}"
    return r


def fempic_params(path, g, m, n_part, fill, generate=True):
    dh = bool(int(g["meta_dh"]))
    b0, b1 = g["bmin"], g["bmax"]
    lines = [
        f"INT nnodes = {m['n_nodes']}", f"INT ncells = {m['n_cells']}", f"INT nparts = {n_part}",
        f"INT num_steps = {int(g['meta_steps'])}", f"REAL dt = {float(g['meta_dt'])!r}",
        f"REAL mass = {float(g['meta_mass'])!r}", f"REAL charge = {float(g['meta_charge'])!r}",
        f"REAL grid_spacing = {float(g['meta_gs'])!r}",
        f"REAL bminx = {float(b0[0])!r}", f"REAL bminy = {float(b0[1])!r}", f"REAL bminz = {float(b0[2])!r}",
        f"REAL bmaxx = {float(b1[0])!r}", f"REAL bmaxy = {float(b1[1])!r}", f"REAL bmaxz = {float(b1[2])!r}",
        f"BOOL opp_global_move = {'true' if dh else 'false'}",
        f"BOOL opp_dh_data_generate = {'true' if generate else 'false'}",
        f"STRING opp_dh_mesh_folder = {os.path.dirname(golden_path(name_of(g)))}/",
        "REAL opp_allocation_multiple = 1.1", "BOOL opp_profile_all = false",
        f"STRING opp_fill = {fill}", "INT opp_fill_period = 4",
    ]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def charge_close(a, b):
    return np.abs(a - b).max() <= DEPOSIT_RTOL * max(np.abs(b).max(), 1e-300)


def name_of(g):
    return "fempic_" + str(g["meta_case"]) + ".npz"


def _fempic_api_run(g, m, fill, td):
    ind, outd = os.path.join(td, "in_" + fill), os.path.join(td, "out_" + fill)
    os.makedirs(ind)
    os.makedirs(outd)
    for k in ["node_pos", "c2n", "c2c", "c_det", "c_vol"]:
        m[k].tofile(os.path.join(ind, k + ".bin"))
    g["in_c_ef"].tofile(os.path.join(ind, "c_ef.bin"))
    g["in_pos"].tofile(os.path.join(ind, "p_pos.bin"))
    g["in_vel"].tofile(os.path.join(ind, "p_vel.bin"))
    g["in_p2c"].astype(np.int32).tofile(os.path.join(ind, "p2c.bin"))
    par = os.path.join(td, fill + ".param")
    fempic_params(par, g, m, g["in_pos"].shape[0], fill)
    run_app(app("fempic_api_app"), par, ind, outd)
    rd = lambda k, t, w: np.fromfile(os.path.join(outd, k + ".bin"), dtype=t).reshape(-1, w)
    recs = np.hstack([rd("final_p2c", np.int32, 1).astype(np.float64), rd("final_pos", np.float64, 3),
                      rd("final_vel", np.float64, 3), rd("final_lc", np.float64, 4)])
    st = np.loadtxt(os.path.join(outd, "steps.txt"), ndmin=2).astype(np.int64)
    return recs, st


@pytest.mark.parametrize("fill", ["Sort_All", "Sort_Periodic"])
def test_fempic_api_sort_finalize(fill):
    """The sort finalize through the opp_lib API (the p2c dat written from the sort's keys, the other
    dats gathered): every particle's record evolves as under hole fill, so the final records are the
    same multiset bit for bit; Sort_All leaves p2c non-decreasing; counts and hops match the reference."""
    g = np.load(golden_path("fempic_small_dh_omp1.npz"), allow_pickle=False)
    n = int(g["meta_nx"])
    m = oppx.box_mesh(n, n, n)
    with tempfile.TemporaryDirectory() as td:
        base, st0 = _fempic_api_run(g, m, "HoleFill_All", td)
        recs, st = _fempic_api_run(g, m, fill, td)
    ref = g["out_steps_iter_before_after_hops_maxhops"]
    assert np.array_equal(st[:, [0, 1, 2, 3]], ref[:, [0, 1, 2, 4]]) and np.array_equal(st, st0)
    order = lambda a: a[np.lexsort(a.T[::-1])]
    assert recs.shape == base.shape and np.array_equal(order(recs), order(base))
    if fill == "Sort_All":
        assert np.all(np.diff(recs[:, 0]) >= 0)


@pytest.mark.parametrize("name,generate", [("fempic_tiny_seq.npz", True), ("fempic_small_dh_omp1.npz", True),
                                           ("fempic_small_dh_omp1.npz", False)])
def test_fempic_api_app_matches_reference(name, generate):
    """generate=False: the overlay is read from the file the reference build dumped for this grid."""
    g = np.load(golden_path(name), allow_pickle=False)
    n = int(g["meta_nx"])
    m = oppx.box_mesh(n, n, n)
    pos, vel, p2c = g["in_pos"], g["in_vel"], g["in_p2c"]
    with tempfile.TemporaryDirectory() as td:
        ind, outd = os.path.join(td, "in"), os.path.join(td, "out")
        os.makedirs(ind)
        os.makedirs(outd)
        for k in ["node_pos", "c2n", "c2c", "c_det", "c_vol"]:
            m[k].tofile(os.path.join(ind, k + ".bin"))
        g["in_c_ef"].tofile(os.path.join(ind, "c_ef.bin"))
        pos.tofile(os.path.join(ind, "p_pos.bin"))
        vel.tofile(os.path.join(ind, "p_vel.bin"))
        p2c.astype(np.int32).tofile(os.path.join(ind, "p2c.bin"))
        par = os.path.join(td, "case.param")
        fempic_params(par, g, m, pos.shape[0], "HoleFill_All", generate=generate)
        run_app(app("fempic_api_app"), par, ind, outd)
        rd = lambda k, t: np.fromfile(os.path.join(outd, k + ".bin"), dtype=t)
        st = np.loadtxt(os.path.join(outd, "steps.txt"), ndmin=2).astype(np.int64)
        ref = g["out_steps_iter_before_after_hops_maxhops"]
        assert np.array_equal(st[:, [0, 1, 2, 3]], ref[:, [0, 1, 2, 4]])
        if int(g["meta_dh"]):
            assert np.array_equal(rd("overlay", np.int32), g["out_overlay"])
        assert np.array_equal(rd("s1_p2c", np.int32), g["out_s1_p2c"])
        assert np.array_equal(rd("s1_pos", np.float64), g["out_s1_pos"])
        assert np.array_equal(rd("s1_lc", np.float64), g["out_s1_lc"])
        assert charge_close(rd("s1_charge", np.float64), g["out_s1_charge"])
        assert np.array_equal(rd("final_p2c", np.int32), g["out_final_p2c"])
        assert np.array_equal(rd("final_pos", np.float64), g["out_final_pos"])
        assert np.array_equal(rd("final_vel", np.float64), g["out_final_vel"])
        assert np.array_equal(rd("final_lc", np.float64), g["out_final_lc"])
        assert charge_close(rd("final_charge", np.float64), g["out_final_charge"])


@pytest.mark.parametrize("name", ["neso_small.npz", "neso_far.npz"])
def test_neso_api_app_matches_reference(name):
    g = np.load(golden_path(name), allow_pickle=False)
    nx, cw, dt = int(g["meta_nx"]), float(g["meta_cw"]), float(g["meta_dt"])
    m = oppx.neso_mesh(nx, nx, cw)
    with tempfile.TemporaryDirectory() as td:
        ind, outd = os.path.join(td, "in"), os.path.join(td, "out")
        os.makedirs(ind)
        os.makedirs(outd)
        m["c_pos_ll"].tofile(os.path.join(ind, "c_pos_ll.bin"))
        m["c2c"].tofile(os.path.join(ind, "c2c.bin"))
        g["in_pos"].tofile(os.path.join(ind, "p_pos.bin"))
        g["in_vel"].tofile(os.path.join(ind, "p_vel.bin"))
        g["in_p2c"].astype(np.int32).tofile(os.path.join(ind, "p2c.bin"))
        par = os.path.join(td, "case.param")
        with open(par, "w") as f:
            f.write("\n".join([f"INT nx = {nx}", f"INT ny = {nx}", f"INT nparts = {g['in_pos'].shape[0]}",
                               f"INT num_steps = {int(g['meta_steps'])}", f"REAL dt = {dt!r}",
                               f"REAL cell_width = {cw!r}", "REAL opp_allocation_multiple = 1.1",
                               "BOOL opp_global_move = false", "STRING opp_fill = HoleFill_All",
                               "BOOL opp_profile_all = false"]) + "\n")
        run_app(app("neso_api_app"), par, ind, outd)
        rd = lambda k, t: np.fromfile(os.path.join(outd, k + ".bin"), dtype=t)
        st = np.loadtxt(os.path.join(outd, "steps.txt"), ndmin=2).astype(np.int64)
        ref = g["out_steps_iter_size_hops_maxhops"]
        assert np.array_equal(st, ref[:, [0, 1, 3]])
        assert np.array_equal(rd("final_p2c", np.int32), g["out_final_p2c"])
        assert np.array_equal(rd("final_pos", np.float64), g["out_final_pos"].ravel())
        assert np.array_equal(rd("final_mdir", np.int32), g["out_final_mdir"].ravel())


@pytest.mark.parametrize("source", ["bin", "hdf5"])
@pytest.mark.parametrize("name", ["fempic_loop_small.npz", "fempic_loop_dh.npz", "fempic_loop_mesh12.npz"])
def test_fempic_loop_app_matches_reference(name, source):
    """The main loop minus the field solve through the drop-in (injection, push, move, deposit, node
    charge density, electric field, OPP_MAX globals) against the reference seq build: particles,
    injected records, boundary potential, field and the field maxima bit-exact; charge density and
    its maximum within 1e-12 relative (FP64 atomic deposit). The mesh12 case runs on a 12-tet box read
    through the mesh-file loader, with its node volumes, node types and inlet faces. source "hdf5":
    the mesh and particle sets come from an OP-PIC HDF5 file holding the same arrays, declared as the
    generated HDF5 main declares them (opp_decl_set_hdf5, opp_decl_map_hdf5, opp_decl_dat_hdf5)."""
    from test_oracle_loop import loop_mesh

    g = np.load(golden_path(name), allow_pickle=False)
    m, ef, bmax = loop_mesh(g)
    with tempfile.TemporaryDirectory() as td:
        ind, outd = os.path.join(td, "in"), os.path.join(td, "out")
        os.makedirs(ind)
        os.makedirs(outd)
        for k in ["node_pos", "c2n", "c2c", "c_det", "c_vol"]:
            m[k].tofile(os.path.join(ind, k + ".bin"))
        ef.tofile(os.path.join(ind, "c_ef.bin"))
        for k in ("sd", "n_vol", "n_type", "n_pot", "if2c", "if_u", "if_v", "if_norm", "if_npos", "if_dist", "rand"):
            g[f"in_{k}"].tofile(os.path.join(ind, ("c_sd" if k == "sd" else k) + ".bin"))
        f = lambda k: float(g[f"meta_{k}"])
        lines = [f"INT nnodes = {m['n_nodes']}", f"INT ncells = {m['n_cells']}", f"INT nfaces = {g['in_if2c'].size}",
                 f"INT nrand = {g['in_rand'].shape[0]}", f"INT inject_count = {int(g['meta_inject'])}",
                 f"INT num_steps = {int(g['meta_steps'])}", f"REAL dt = {f('dt')!r}", f"REAL mass = {f('mass')!r}",
                 f"REAL charge = {f('charge')!r}", f"REAL spwt = {f('spwt')!r}",
                 f"REAL ion_velocity = {f('ion_velocity')!r}", f"REAL wall_potential = {f('wall_potential')!r}",
                 f"REAL grid_spacing = {f('gs')!r}", "REAL bminx = 0.0", "REAL bminy = 0.0", "REAL bminz = 0.0",
                 *[f"REAL bmax{a} = {v!r}" for a, v in zip("xyz", bmax)],
                 f"BOOL opp_global_move = {'true' if int(g['meta_dh']) else 'false'}",
                 "BOOL opp_dh_data_generate = true", "REAL opp_allocation_multiple = 1.1",
                 "BOOL opp_profile_all = false", "STRING opp_fill = HoleFill_All", "INT opp_fill_period = 4"]
        if source == "hdf5":
            import h5write

            nc, nn, nf = m["n_cells"], m["n_nodes"], g["in_if2c"].size
            one = lambda v: np.array([v], np.int32)
            col = lambda a, d: np.ascontiguousarray(a).reshape(-1, d)
            h5 = os.path.join(td, "mesh.h5")
            h5write.write(h5, {
                "mesh_nodes": one(nn), "mesh_cells": one(nc), "inlet_faces_cells": one(nf), "particles": one(0),
                "dummy particles": one(0), "c_v_n_map": col(m["c2n"], 4), "c_v_c_map": col(m["c2c"], 4),
                "if_v_c_map": col(g["in_if2c"], 1), "c_det": col(m["c_det"], 16), "c_volume": col(m["c_vol"], 1),
                "c_ef": col(ef, 3), "c_shape_deri": col(g["in_sd"], 12), "c_gbl_id": col(np.arange(nc, dtype=np.int32), 1),
                "n_vol": col(g["in_n_vol"], 1), "n_potential": col(g["in_n_pot"], 1),
                "n_charge_den": np.zeros((nn, 1)), "n_pos": col(m["node_pos"], 3), "n_type": col(g["in_n_type"], 1),
                "n_bnd_pot": np.zeros((nn, 1)), "iface_v_norm": col(g["in_if_v"], 3),
                "iface_u_norm": col(g["in_if_u"], 3), "iface_norm": col(g["in_if_norm"], 3),
                "iface_dist": col(g["in_if_dist"], 1), "iface_n_pos": col(g["in_if_npos"], 9)})
            lines.append(f"STRING hdf_filename = {h5}")
        par = os.path.join(td, "case.param")
        with open(par, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        run_app(app("fempic_loop_app"), par, ind, outd)
        rd = lambda k, t: np.fromfile(os.path.join(outd, k + ".bin"), dtype=t)
        assert np.array_equal(rd("bnd_pot", np.float64), g["out_bnd_pot"])
        assert np.array_equal(rd("dist1_p2c", np.int32), g["out_dist1_p2c"])
        for k, t in (("pos", np.float64), ("vel", np.float64), ("p2c", np.int32)):
            assert np.array_equal(rd("inj1_" + k, t), g["out_inj1_" + k]), "injected " + k
        rows = [ln.split() for ln in open(os.path.join(outd, "steps.txt")).read().splitlines() if ln.strip()]
        st = np.array([[int(x) for x in r[:4]] for r in rows], np.int64)
        mx = np.array([[float.fromhex(x) for x in r[4:7]] for r in rows])
        ref, rmx = g["out_steps_iter_before_after_hops_maxhops"], g["out_steps_max_cef_chg_pot"]
        assert np.array_equal(st, ref[:, [0, 1, 2, 4]])
        assert np.array_equal(mx[:, 0], rmx[:, 0]) and np.array_equal(mx[:, 2], rmx[:, 2])
        assert np.all(np.abs(mx[:, 1] - rmx[:, 1]) <= DEPOSIT_RTOL * np.abs(rmx[:, 1]))
        for t in ("s1", "final"):
            for k, dt_ in (("pos", np.float64), ("vel", np.float64), ("lc", np.float64), ("p2c", np.int32),
                           ("ef", np.float64)):
                assert np.array_equal(rd(f"{t}_{k}", dt_), g[f"out_{t}_{k}"]), f"{t} {k}"
            assert charge_close(rd(f"{t}_charge", np.float64), g[f"out_{t}_charge"]), f"{t} charge"


@pytest.mark.parametrize("name", ["cabana_small.npz", "cabana_grid16.npz"])
def test_cabana_api_app_matches_reference(name):
    """CabanaPIC's move_deposit through the drop-in: particles bit-exact (digests for the larger
    case), accumulators within 1e-12 relative, per-step max hops as the reference's."""
    import hashlib

    from test_oracle_cabana import cabana_case

    g = np.load(golden_path(name), allow_pickle=False)
    d = cabana_case(g)
    with tempfile.TemporaryDirectory() as td:
        ind, outd = os.path.join(td, "in"), os.path.join(td, "out")
        os.makedirs(ind)
        os.makedirs(outd)
        for k, fn in (("c2c6", "c2ngc"), ("interp", "c_interp"), ("pos", "p_pos"), ("vel", "p_vel"),
                      ("weight", "p_weight"), ("p2c", "p2c")):
            np.ascontiguousarray(d[k]).tofile(os.path.join(ind, fn + ".bin"))
        cd = g["meta_cdt_d"]
        lines = [f"INT ncells = {d['n_cells']}", f"INT nparts = {d['p2c'].size}", f"INT num_steps = {int(g['meta_steps'])}",
                 f"REAL qsp = {float(g['meta_qsp'])!r}", f"REAL qdt_half_mc = {float(g['meta_qdt_2mc'])!r}",
                 f"REAL cdt_dx = {float(cd[0])!r}", f"REAL cdt_dy = {float(cd[1])!r}", f"REAL cdt_dz = {float(cd[2])!r}",
                 "REAL opp_allocation_multiple = 1.1", "BOOL opp_profile_all = false", "STRING opp_fill = HoleFill_All",
                 "INT opp_fill_period = 4", "BOOL opp_global_move = false"]
        par = os.path.join(td, "case.param")
        with open(par, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        run_app(app("cabana_api_app"), par, ind, outd)
        rd = lambda k, t: np.fromfile(os.path.join(outd, k + ".bin"), dtype=t)
        st = np.loadtxt(os.path.join(outd, "steps.txt"), ndmin=2).astype(np.int64)
        ref = g["out_steps_iter_size_hops_maxhops"]
        assert np.array_equal(st, ref[:, [0, 1, 3]])
        for t in ("s1", "final"):
            r = g[f"out_{t}_acc"]
            assert np.abs(rd(f"{t}_acc", np.float64) - r).max() <= DEPOSIT_RTOL * np.abs(r).max(), f"{t} acc"
            for k, dt_ in (("pos", np.float64), ("vel", np.float64), ("streak", np.float64), ("p2c", np.int32)):
                a = rd(f"{t}_{k}", dt_)
                if f"out_{t}_{k}" in g:
                    assert np.array_equal(a, g[f"out_{t}_{k}"]), f"{t} {k}"
                else:
                    assert hashlib.sha256(a.tobytes()).hexdigest() == str(g[f"sha_{t}_{k}"]), f"{t} {k}"


def test_generated_neso_main_unchanged_matches_reference():
    """OP-PIC's own generated NESO-advection main (advec_opp.cpp, compiled unchanged against the drop-in's
    headers and linked with libopp_mi355x.so) on the deck it was fixtured with: the per-step verify error
    count and particle count it logs, and the final particle dats (dumped at opp_exit), equal the
    reference seq build's bit for bit. The deck's particles come from the main's own std::mt19937 seeds.
    Not compared: the log's "Hops: Max" field, which the reference never assigns (it stays -1) and the
    drop-in fills with the move's longest walk."""
    binary = os.path.join(API_DIR, "advec_opp_dropin")
    if not os.path.exists(binary):
        pytest.skip("advec_opp_dropin is built only where the OP-PIC source tree is present")
    from make_neso_app_golden import parse_log, read_dump

    g = np.load(golden_path("neso_app_small.npz"), allow_pickle=False)
    with tempfile.TemporaryDirectory() as td:
        print(f"[api] advec_opp_dropin {golden_path('neso_app_small.param')}", flush=True)
        r = subprocess.run([binary, golden_path("neso_app_small.param")], capture_output=True, text=True, timeout=300,
                           cwd=td, env=dict(os.environ, OPPX_APP_DUMP=td))
        assert r.returncode == 0, f"rc={r.returncode}\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}"
        steps = parse_log(r.stdout)
        out = read_dump(td, int(steps[-1, 2]))
    assert np.array_equal(steps, g["steps"])
    for k in ["pos", "vel", "mdir", "p2c", "index"]:
        assert np.array_equal(out[k], g[k]), k
