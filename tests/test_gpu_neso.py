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


@pytest.mark.parametrize("n,seed", [(0, 1), (1, 2), (5000, 3), (200_003, 4)])
def test_gpu_neso_verify_counts_misplaced(n, seed):
    """oppx_neso_verify (the verify loop of apps/app_neso_advection/kernels.h:115-150) on particles of
    which some sit outside their cell, some exactly on a cell edge (inside: the test is inclusive), and
    some hold a cell whose global index is negative; the count equals the closed form computed here."""
    import torch

    nx, ny, cw = 37, 23, 0.25
    rng = np.random.default_rng(seed)
    n_cells = nx * ny
    gidx = np.arange(n_cells, dtype=np.int32)
    gidx[rng.random(n_cells) < 0.05] = -3  # a corrupt global index: ix, iy negative
    p2c = rng.integers(0, n_cells, n).astype(np.int32)
    g = gidx[p2c].astype(np.int64)
    ix, iy = g % nx, g // nx
    # C division: (-3) / nx == 0 and -3 - 0 == -3, so ix < 0 for the corrupt entries
    ix = np.where(g < 0, g - np.trunc(g / nx).astype(np.int64) * nx, ix)
    iy = np.where(g < 0, np.trunc(g / nx).astype(np.int64), iy)
    u = rng.random((n, 2)) * 1.2 - 0.1  # about a third of the particles leave their box
    edge = rng.random(n) < 0.05
    u[edge, 0] = 1.0  # exactly on the upper x edge: inside
    pos = np.stack([(ix * cw) + u[:, 0] * cw, (iy * cw) + u[:, 1] * cw], axis=1)
    pos[edge, 0] = ix[edge] * cw + cw
    llx, lly = ix * cw, iy * cw
    outside = (g < 0) | (pos[:, 0] < llx) | (pos[:, 0] > llx + cw) | (pos[:, 1] < lly) | (pos[:, 1] > lly + cw)
    expect = int(outside.sum())
    cap = n + 5
    pos_d = torch.zeros((2, cap), dtype=torch.float64, device="cuda")
    pos_d[:, :n] = torch.from_numpy(pos.T.copy())
    p2c_d = torch.zeros(cap, dtype=torch.int32, device="cuda")
    p2c_d[:n] = torch.from_numpy(p2c)
    gidx_d = torch.from_numpy(gidx).to("cuda")
    ctr = torch.zeros(oppx.COUNTER_WORDS, dtype=torch.int32, device="cuda")
    ctr[0] = n
    count = torch.full((1,), 7, dtype=torch.int32, device="cuda")  # the kernel adds to what is there
    oppx.check(oppx.lib().oppx_neso_verify(n, cap, pos_d.data_ptr(), p2c_d.data_ptr(), gidx_d.data_ptr(), n_cells, nx,
                                           cw, ctr.data_ptr(), count.data_ptr(),
                                           torch.cuda.current_stream().cuda_stream), "neso_verify")
    assert int(count.item()) == 7 + expect
    assert int(ctr[6].item()) == 0  # faults
