"""GPU parity of the NESO-advection hot path (periodic drift + 2-D multi-hop walk) against the
reference seq build's fixtures: positions, cells, move directions and hop counts bit-exact."""

import hashlib

import numpy as np
import pytest

import oppx
from conftest import golden_path

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("grid", [True, False], ids=["grid-records", "gathered-records"])
@pytest.mark.parametrize("name", ["neso_small.npz", "neso_far.npz", "neso_c3_mini.npz"])
def test_gpu_neso_matches_reference(name, grid):
    # both walks (records computed on the verified periodic grid, or gathered) give the reference's
    # positions, cells, directions and hop counts
    g = np.load(golden_path(name), allow_pickle=False)
    nx, cw = int(g["meta_nx"]), float(g["meta_cw"])
    m = oppx.neso_mesh(nx, nx, cw)
    pos, vel, p2c = oppx.neso_particles(int(g["meta_n_part"]), int(g["meta_seed"]), m, float(g["meta_vmax"]))
    gpu = oppx.NesoDevice(m, pos, vel, p2c, grid=grid)
    assert gpu.grid == grid
    dt = float(g["meta_dt"])
    st = g["out_steps_iter_size_hops_maxhops"]
    for s in range(int(g["meta_steps"])):
        gpu.step(dt, record_hops=(s == 0))
        c = gpu.read_counters()
        assert c["faults"] == 0
        assert (c["total_hops"], c["max_hops"]) == (st[s, 2], st[s, 3]), f"step {s}: {c}"
        if s == 0 and "out_s1_p2c" in g:
            o = gpu.download()
            assert np.array_equal(o["p2c"], g["out_s1_p2c"]) and np.array_equal(o["pos"].ravel(), g["out_s1_pos"])
            assert np.array_equal(o["mdir"].ravel(), g["out_s1_mdir"]) and np.array_equal(o["hops"], g["out_s1_hops"])
    o = gpu.download()
    if "out_final_p2c" in g:
        assert np.array_equal(o["p2c"], g["out_final_p2c"]) and np.array_equal(o["pos"].ravel(), g["out_final_pos"])
        assert np.array_equal(o["mdir"].ravel(), g["out_final_mdir"])
    else:
        assert sha(o["p2c"]) == str(g["sha_final_p2c"]) and sha(o["pos"]) == str(g["sha_final_pos"])
        assert sha(o["mdir"]) == str(g["sha_final_mdir"])


def rows(o):
    r = np.concatenate([o["p2c"][:, None].astype(np.int64), o["pos"].view(np.int64),
                        o["mdir"].astype(np.int64)], axis=1)
    return r[np.lexsort(r.T[::-1])]


@pytest.mark.parametrize("name", ["neso_small.npz", "neso_c3_mini.npz"])
def test_gpu_neso_periodic_sort_only_permutes(name):
    # sorting by cell every other step reorders the particles and changes nothing else: per-step
    # hop totals and maxima stay the reference's, and the final particle records are the same rows
    g = np.load(golden_path(name), allow_pickle=False)
    nx, cw = int(g["meta_nx"]), float(g["meta_cw"])
    m = oppx.neso_mesh(nx, nx, cw)
    pos, vel, p2c = oppx.neso_particles(int(g["meta_n_part"]), int(g["meta_seed"]), m, float(g["meta_vmax"]))
    plain, srt = oppx.NesoDevice(m, pos, vel, p2c), oppx.NesoDevice(m, pos, vel, p2c)
    dt = float(g["meta_dt"])
    st = g["out_steps_iter_size_hops_maxhops"]
    srt.sort()
    for s in range(int(g["meta_steps"])):
        plain.step(dt)
        srt.step(dt)
        c = srt.read_counters()
        assert c["faults"] == 0 and (c["total_hops"], c["max_hops"]) == (st[s, 2], st[s, 3]), f"step {s}: {c}"
        if s % 2 == 1:
            srt.sort()
            cells = srt.p2c[: srt.n].cpu().numpy()
            assert (np.diff(cells) >= 0).all()
    assert np.array_equal(rows(plain.download()), rows(srt.download()))


def test_gpu_quad_grid_check_rejects_other_meshes():
    # one corner moved by one ulp, or one neighbour swapped: the mesh is not the grid, and the
    # gathered walk runs
    nx, cw = 16, 0.5
    m = oppx.neso_mesh(nx, nx, cw)
    pos, vel, p2c = oppx.neso_particles(1000, 7, m, 1.0)
    assert oppx.NesoDevice(m, pos, vel, p2c).grid
    m1 = dict(m, c_pos_ll=m["c_pos_ll"].copy())
    m1["c_pos_ll"].reshape(-1)[2 * 37] = np.nextafter(m1["c_pos_ll"].reshape(-1)[2 * 37], 1e9)
    assert not oppx.NesoDevice(m1, pos, vel, p2c).grid
    m2 = dict(m, c2c=m["c2c"].copy())
    c = m2["c2c"].reshape(-1, 4)
    c[5, 0], c[5, 1] = c[5, 1], c[5, 0]
    assert not oppx.NesoDevice(m2, pos, vel, p2c).grid
