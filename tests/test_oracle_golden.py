"""The oracle against the reference build's golden fixtures (tests/golden, written by
make_fempic_golden.py from the reference seq/omp libraries) and against closed forms.

CPU only. These pin the oracle before the GPU path is compared with it.
"""

import hashlib

import numpy as np
import pytest

import oppx
import oracle
from conftest import golden_path


def load(name):
    return np.load(golden_path(name), allow_pickle=False)


def mesh_of(g):
    n = int(g["meta_nx"])
    if str(g["meta_case"]).startswith("mesh12"):  # 12-tet box read through the mesh-file loader
        nz = int(g["meta_nz"])
        return oppx.box12_mesh(n, int(g["meta_ny"]), nz, 1.0, 1.0, nz / n)
    return oppx.box_mesh(n, n, n)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def consts(g):
    return float(g["meta_charge"]), float(g["meta_mass"]), float(g["meta_dt"])


def inputs_of(g, m):
    if "in_pos" in g:
        return g["in_c_ef"], g["in_pos"], g["in_vel"], g["in_p2c"]
    ef = oppx.box_field_sines(m)
    pos, vel, p2c = oppx.box_particles(int(g["meta_n_part"]), int(g["meta_seed"]), float(g["meta_sigma"]), m)
    return ef, pos, vel, p2c


def test_inputs_are_host_independent():
    # the builders must reproduce the exact inputs the reference consumed
    for name in ["fempic_tiny_seq.npz", "fempic_small_dh_omp1.npz"]:
        g = load(name)
        m = mesh_of(g)
        ef = oppx.box_field_sines(m)
        pos, vel, p2c = oppx.box_particles(int(g["meta_n_part"]), int(g["meta_seed"]), float(g["meta_sigma"]), m)
        assert np.array_equal(ef, g["in_c_ef"])
        assert np.array_equal(pos, g["in_pos"]) and np.array_equal(vel, g["in_vel"])
        assert np.array_equal(p2c, g["in_p2c"])
    g = load("fempic_c1.npz")
    m = mesh_of(g)
    pos, vel, p2c = oppx.box_particles(int(g["meta_n_part"]), int(g["meta_seed"]), float(g["meta_sigma"]), m)
    assert sha(pos) == str(g["sha_in_pos"]) and sha(vel) == str(g["sha_in_vel"]) and sha(p2c) == str(g["sha_in_p2c"])


@pytest.mark.parametrize("name,threads", [("fempic_tiny_seq.npz", 1), ("fempic_small_dh_omp1.npz", 1),
                                          ("fempic_small_dh_omp4.npz", 4), ("fempic_mesh12_dh.npz", 1)])
def test_oracle_matches_reference_bit_exact(name, threads):
    g = load(name)
    m = mesh_of(g)
    ef, pos, vel, p2c = inputs_of(g, m)
    kw = {}
    if int(g["meta_dh"]):
        kw = dict(overlay=g["out_overlay"], bmin=g["bmin"], bmax=g["bmax"], gs=float(g["meta_gs"]))
    st = oracle.FempicState(m, ef, pos, vel, p2c, **kw)
    q, ms, dt = consts(g)
    steps = g["out_steps_iter_before_after_hops_maxhops"]
    for s in range(int(g["meta_steps"])):
        r = st.step(q, ms, dt, threads=threads)
        assert (r["size"], r["total_hops"], r["max_hops"]) == tuple(steps[s, 2:5]), f"step {s}"
        if s == 0:
            L = st.live()
            assert np.array_equal(L["p2c"], g["out_s1_p2c"])
            assert np.array_equal(L["pos"].ravel(), g["out_s1_pos"])
            assert np.array_equal(L["vel"].ravel(), g["out_s1_vel"])
            assert np.array_equal(L["lc"].ravel(), g["out_s1_lc"])
            assert np.array_equal(L["charge"], g["out_s1_charge"])
    L = st.live()
    assert np.array_equal(L["p2c"], g["out_final_p2c"])
    assert np.array_equal(L["pos"].ravel(), g["out_final_pos"])
    assert np.array_equal(L["lc"].ravel(), g["out_final_lc"])
    assert np.array_equal(L["charge"], g["out_final_charge"])


def test_oracle_per_particle_hops_match_reference():
    g = load("fempic_small_dh_omp1.npz")
    m = mesh_of(g)
    ef, pos, vel, p2c = inputs_of(g, m)
    st = oracle.FempicState(m, ef, pos, vel, p2c, overlay=g["out_overlay"], bmin=g["bmin"], bmax=g["bmax"],
                            gs=float(g["meta_gs"]))
    q, ms, dt = consts(g)
    st.push(q, ms, dt)
    removed, hops = st.move()
    assert np.array_equal(hops, g["out_s1_hops"])


def test_oracle_c1_scale_matches_reference():
    # 48,000 tets, 1M particles, 100 steps, 33^3 overlay, deposit with the reference's 8-thread order
    g = load("fempic_c1.npz")
    m = mesh_of(g)
    ef, pos, vel, p2c = inputs_of(g, m)
    st = oracle.FempicState(m, ef, pos, vel, p2c, overlay=g["out_overlay"], bmin=g["bmin"], bmax=g["bmax"],
                            gs=float(g["meta_gs"]))
    q, ms, dt = consts(g)
    steps = g["out_steps_iter_before_after_hops_maxhops"]
    for s in range(int(g["meta_steps"])):
        r = st.step(q, ms, dt, threads=8)
        assert (r["size"], r["total_hops"], r["max_hops"]) == tuple(steps[s, 2:5]), f"step {s}"
        if s == 0:
            assert sha(st.live()["p2c"]) == str(g["sha_s1_p2c"])
            assert np.array_equal(st.live()["charge"], g["out_s1_charge"])
    L = st.live()
    assert sha(L["p2c"]) == str(g["sha_final_p2c"])
    assert sha(L["pos"]) == str(g["sha_final_pos"])
    assert np.array_equal(L["charge"], g["out_final_charge"])


@pytest.mark.parametrize("name", ["fempic_small_dh_omp1.npz", "fempic_c1.npz"])
def test_overlay_builders_match_reference(name):
    g = load(name)
    m = mesh_of(g)
    gs = float(g["meta_gs"])
    ref = g["out_overlay"]
    assert np.array_equal(oppx.dh_dims(g["bmin"], g["bmax"], gs), g["out_overlay_dims"])
    assert np.array_equal(oppx.dh_overlay(m, g["bmin"], g["bmax"], gs), ref)  # product builder
    assert np.array_equal(oppx.dh_overlay_geom(m, g["bmin"], g["bmax"], gs), ref)  # from c_det/c_volume only
    if m["n_cells"] <= 5000:
        assert np.array_equal(oracle.dh_overlay_bruteforce(m, g["bmin"], g["bmax"], gs), ref)


# ---------------------------------------------------------------------------------------------
# closed forms that pin what the fixtures leave open
# ---------------------------------------------------------------------------------------------

def test_barycentric_closed_form():
    # lc_i = 1 at vertex i and 0 at the others; sum of lc = 1 at any point
    m = oppx.box_mesh(3, 2, 4, 1.0, 2.0, 0.5)
    for c in [0, 5, 17, m["n_cells"] - 1]:
        det, vol = m["c_det"][c].reshape(4, 4), m["c_vol"][c]
        for v in range(4):
            x, y, z = m["node_pos"][m["c2n"][c, v]]
            lc = (det[:, 0] - det[:, 1] * x + det[:, 2] * y - det[:, 3] * z) / (6 * vol)
            assert np.allclose(lc, np.eye(4)[v], atol=1e-12)
    # all volumes positive and summing to the box volume
    assert (m["c_vol"] > 0).all() and abs(m["c_vol"].sum() - 1.0) < 1e-12


def test_face_neighbours_closed_form():
    m = oppx.box_mesh(3, 3, 3)
    c2n, c2c = m["c2n"], m["c2c"]
    for c in range(m["n_cells"]):
        for v in range(4):
            nb = c2c[c, v]
            face = set(c2n[c]) - {c2n[c, v]}
            if nb < 0:  # boundary face: all three nodes on the box surface, on one common plane
                pts = m["node_pos"][list(face)]
                assert any(np.all(pts[:, d] == 0.0) or np.all(pts[:, d] == 1.0) for d in range(3))
            else:
                assert face <= set(c2n[nb]) and nb != c
                assert c in c2c[nb]


def test_kuhn_containment_closed_form():
    m = oppx.box_mesh(5, 5, 5)
    pos, vel, p2c = oppx.box_particles(20000, 99, 0.01, m)
    det, vol = m["c_det"][p2c].reshape(-1, 4, 4), m["c_vol"][p2c]
    x, y, z = pos[:, 0:1], pos[:, 1:2], pos[:, 2:3]
    lc = (det[:, :, 0] - det[:, :, 1] * x + det[:, :, 2] * y - det[:, :, 3] * z) / (6 * vol[:, None])
    assert (lc > -1e-10).all() and (lc < 1 + 1e-10).all()


def holefill_closed_form(marks, removed):
    """k-th hole (ascending) below S = N - R receives the k-th live record (descending) at or above S."""
    n = marks.size
    s = n - removed
    holes = np.nonzero(marks[:s] == oracle.MAX_CELL_INDEX)[0]
    lives = np.nonzero(marks[s:] != oracle.MAX_CELL_INDEX)[0][::-1] + s
    assert holes.size == lives.size
    perm = np.arange(s)
    perm[holes] = lives
    return perm


@pytest.mark.parametrize("n,frac,seed", [(1000, 0.1, 1), (4096, 0.5, 2), (777, 0.97, 3), (50, 0.0, 4), (64, 1.0, 5)])
def test_holefill_closed_form(n, frac, seed):
    rng = np.random.default_rng(seed)
    marks = np.where(rng.random(n) < frac, oracle.MAX_CELL_INDEX, rng.integers(0, 100, n)).astype(np.int32)
    removed = int((marks == oracle.MAX_CELL_INDEX).sum())
    ids = np.arange(n, dtype=np.int64).astype(np.float64).reshape(n, 1).copy()
    p2c = marks.copy()
    ptrs = (oracle.c_void_p * 2)(ids.ctypes.data, p2c.ctypes.data)
    new_n = oracle.lib().orc_holefill(n, removed, oracle._p(marks), ptrs, oracle._p(np.array([1, 1], np.int32)),
                                      oracle._p(np.array([8, 4], np.int32)), 2)
    assert new_n == n - removed
    expect = holefill_closed_form(marks, removed)
    assert np.array_equal(ids[:new_n, 0].astype(np.int64), expect)
    assert (p2c[:new_n] != oracle.MAX_CELL_INDEX).all()


def test_deposit_conserves_charge():
    g = load("fempic_small_dh_omp1.npz")
    m = mesh_of(g)
    ef, pos, vel, p2c = inputs_of(g, m)
    st = oracle.FempicState(m, ef, pos, vel, p2c, overlay=g["out_overlay"], bmin=g["bmin"], bmax=g["bmax"],
                            gs=float(g["meta_gs"]))
    q, ms, dt = consts(g)
    st.step(q, ms, dt)
    # every live particle deposits its 4 weights, which sum to 1
    assert abs(st.charge.sum() - st.n) < 1e-9 * max(st.n, 1)
    assert abs(st.lc[: st.n].sum() - st.n) < 1e-9 * max(st.n, 1)


def test_sort_by_cell_contract():
    rng = np.random.default_rng(7)
    n = 5000
    p2c = rng.integers(0, 300, n).astype(np.int32)
    ids = np.arange(n, dtype=np.float64).reshape(n, 1).copy()
    keys = p2c.copy()
    perm = oracle.sort_by_cell(n, [ids, keys], p2c)
    assert np.array_equal(perm, np.argsort(p2c, kind="stable"))
    assert np.all(np.diff(keys) >= 0)
    assert np.array_equal(ids[:, 0].astype(np.int64), perm)


def test_morton_numbered_mesh_is_the_same_geometry():
    # Morton numbering permutes cubes and nodes only: the walk over it gives the same positions,
    # lc and hop counts, and cells that map to the x-ordered ones through the cube permutation
    nx = 5
    a, b = oppx.box_mesh(nx, nx, nx), oppx.box_mesh(nx, nx, nx, order="morton")
    cent = lambda m: np.round(m["node_pos"][m["c2n"]].mean(axis=1) * 4 * nx).astype(np.int64)
    ka = {tuple(r): i for i, r in enumerate(cent(a))}
    to_a = np.array([ka[tuple(r)] for r in cent(b)])  # morton cell -> x-ordered cell
    assert np.array_equal(np.sort(to_a), np.arange(a["n_cells"]))
    assert np.array_equal(a["c_vol"][to_a], b["c_vol"])
    c2c_b = np.where(b["c2c"] >= 0, to_a[b["c2c"]], -1)
    assert np.array_equal(np.sort(c2c_b, axis=1), np.sort(a["c2c"][to_a], axis=1))
    ef_a, ef_b = oppx.box_field_sines(a), oppx.box_field_sines(b)
    pa, va, ca = oppx.box_particles(3000, 9, 0.02, a)
    pb, vb, cb = oppx.box_particles(3000, 9, 0.02, b)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb) and np.array_equal(to_a[cb], ca)
    sa = oracle.FempicState(a, ef_a, pa, va, ca)
    sb = oracle.FempicState(b, ef_b, pb, vb, cb)
    for _ in range(4):
        sa.step(1e-3, 1.0, 0.02)
        sb.step(1e-3, 1.0, 0.02)
    ea, eb = sa.live(), sb.live()
    assert ea["n"] == eb["n"]
    for k in ("pos", "vel", "lc"):
        assert np.array_equal(ea[k], eb[k]), k
    assert np.array_equal(to_a[eb["p2c"]], ea["p2c"])
