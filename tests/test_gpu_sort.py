"""GPU particle sort against the oracle: stable by (p2c, index), removed particles last and
dropped by the finalize; the Sort_All step policy bit-exact against the oracle's."""

import math

import numpy as np
import pytest
import torch

import oppx
import oracle
from conftest import golden_path

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,n_cells,frac_removed,seed,items", [(1, 5, 0.0, 0, 16), (1000, 48, 0.1, 1, 4),
                                                               (70001, 300, 0.3, 2, 8), (300000, 6_000_000, 0.05, 3, 16),
                                                               (300000, 6_000_000, 0.05, 3, 4), (5000, 7, 0.0, 4, 8),
                                                               (4096, 100000, 1.0, 5, 16), (20000, 40_000_000, 0.1, 6, 8),
                                                               (300000, 6_000_000, 0.05, 3, 32), (70001, 300, 0.3, 2, 32),
                                                               (8191, 7, 0.0, 7, 32)])
def test_gpu_sort_matches_oracle(n, n_cells, frac_removed, seed, items):
    saved = oppx.get_tuning()
    oppx.set_tuning(**dict(saved, sort_items=items))
    try:
        _sort_case(n, n_cells, frac_removed, seed)
    finally:
        oppx.set_tuning(**saved)


@pytest.mark.parametrize("n,n_cells,frac_removed,seed", [(1, 5, 0.0, 0), (70001, 300, 0.3, 2),
                                                         (300000, 6_000_000, 0.05, 3), (5000, 7, 0.0, 4),
                                                         (2_000_000, 6_000_000, 0.02, 7)])
def test_gpu_sort_permutation_only(n, n_cells, frac_removed, seed):
    """n_dats = 0 with perm_out (the lazy policy's sort): the last digit pass writes the permutation
    itself, which must be the stable order of the keys."""
    rng = np.random.default_rng(seed)
    p2c = rng.integers(0, n_cells, n).astype(np.int32)
    p2c[rng.random(n) < frac_removed] = oppx.MAX_CELL_INDEX
    if seed == 4:
        p2c[:] = 3
    if seed == 7:  # nearly sorted, as after a move of cell-ordered particles
        p2c = np.sort(p2c)
        swap = rng.integers(0, n - 1, n // 50)
        p2c[swap], p2c[swap + 1] = p2c[swap + 1], p2c[swap].copy()
    L = oppx.lib()
    cap = n + 5
    k_in = torch.zeros(cap, dtype=torch.int32, device="cuda")
    k_in[:n] = torch.from_numpy(p2c)
    perm = torch.full((cap,), -7, dtype=torch.int32, device="cuda")
    ctr = torch.zeros(oppx.COUNTER_WORDS, dtype=torch.int32, device="cuda")
    ctr[0] = n
    sb = int(L.oppx_sort_scratch_bytes(n))
    scratch = torch.zeros(sb, dtype=torch.uint8, device="cuda")
    oppx.check(L.oppx_sort_by_cell(n, cap, ctr.data_ptr(), n_cells, k_in.data_ptr(), None, None, 0, perm.data_ptr(), 0,
                                   scratch.data_ptr(), sb, torch.cuda.current_stream().cuda_stream), "sort")
    got = perm.cpu().numpy()
    assert np.array_equal(got[:n], np.argsort(p2c, kind="stable"))
    assert (got[n:] == -7).all()  # nothing written past the live range
    c = ctr.cpu().numpy()
    assert c[0] == n and c[1] == int((p2c == oppx.MAX_CELL_INDEX).sum())  # drop_removed = 0 keeps the count


def _sort_case(n, n_cells, frac_removed, seed):
    rng = np.random.default_rng(seed)
    p2c = rng.integers(0, n_cells, n).astype(np.int32)
    p2c[rng.random(n) < frac_removed] = oppx.MAX_CELL_INDEX
    if seed == 4:
        p2c[:] = 3  # one long run of equal keys: order must be the index order
    vals = rng.random((n, 3))
    cap = n + 5
    L = oppx.lib()
    dev = "cuda"
    v_in = torch.zeros((3, cap), dtype=torch.float64, device=dev)
    v_in[:, :n] = torch.from_numpy(vals.T.copy())
    k_in = torch.zeros(cap, dtype=torch.int32, device=dev)
    k_in[:n] = torch.from_numpy(p2c)
    v_out, k_out = torch.zeros_like(v_in), torch.zeros_like(k_in)
    perm = torch.zeros(cap, dtype=torch.int32, device=dev)
    ctr = torch.zeros(oppx.COUNTER_WORDS, dtype=torch.int32, device=dev)
    ctr[0] = n
    sb = int(L.oppx_sort_scratch_bytes(n))
    scratch = torch.zeros(sb, dtype=torch.uint8, device=dev)
    ins = (oppx.DatDesc * 2)(oppx.DatDesc(v_in.data_ptr(), 3, 8), oppx.DatDesc(k_in.data_ptr(), 1, 4))
    outs = (oppx.DatDesc * 2)(oppx.DatDesc(v_out.data_ptr(), 3, 8), oppx.DatDesc(k_out.data_ptr(), 1, 4))
    oppx.check(L.oppx_sort_by_cell(n, cap, ctr.data_ptr(), n_cells, k_in.data_ptr(), ins, outs, 2, perm.data_ptr(), 1,
                                   scratch.data_ptr(), sb, torch.cuda.current_stream().cuda_stream), "sort")
    removed = int((p2c == oppx.MAX_CELL_INDEX).sum())
    c = ctr.cpu().numpy()
    assert c[0] == n - removed and c[1] == removed
    expect = np.argsort(p2c, kind="stable")  # MAX sorts last
    got = perm[:n].cpu().numpy()
    live = n - removed
    assert np.array_equal(got[:live], expect[:live])
    assert np.array_equal(np.sort(got), np.arange(n))  # a permutation, removed ones in the tail
    assert np.array_equal(k_out[:n].cpu().numpy(), p2c[got])
    assert np.array_equal(v_out[:, :n].cpu().numpy().T, vals[got])
    # the oracle's sort agrees on the live range
    arrs = [vals.copy(), p2c.copy()]
    operm = oracle.sort_by_cell(n, arrs, p2c)
    assert np.array_equal(operm[:live], got[:live])


def test_gpu_sort_fill_policy_steps_match_oracle():
    g = np.load(golden_path("fempic_small_dh_omp1.npz"), allow_pickle=False)
    n = int(g["meta_nx"])
    m = oppx.box_mesh(n, n, n)
    kw = dict(overlay=g["out_overlay"], bmin=tuple(g["bmin"]), bmax=tuple(g["bmax"]), gs=float(g["meta_gs"]))
    gpu = oppx.FempicDevice(m, g["in_c_ef"], g["in_pos"], g["in_vel"], g["in_p2c"], **kw)
    ref = oracle.FempicState(m, g["in_c_ef"], g["in_pos"], g["in_vel"], g["in_p2c"], **kw)
    q, ms, dt = float(g["meta_charge"]), float(g["meta_mass"]), float(g["meta_dt"])
    steps = g["out_steps_iter_before_after_hops_maxhops"]
    for s in range(int(g["meta_steps"])):
        gpu.step_fused(q, ms, dt, fill="sort")
        r = ref.step_sorted(q, ms, dt)
        c = gpu.read_counters()
        # counts and hops do not depend on the fill policy
        assert (c["size"], c["total_hops"], c["max_hops"]) == tuple(steps[s, 2:5])
        assert (c["size"], c["total_hops"]) == (r["size"], r["total_hops"])
        out, exp = gpu.download(), ref.live()
        assert np.array_equal(out["p2c"], exp["p2c"]) and np.all(np.diff(out["p2c"]) >= 0)
        assert np.array_equal(out["pos"], exp["pos"]) and np.array_equal(out["lc"], exp["lc"])
        scale = max(np.abs(exp["charge"]).max(), 1e-300)
        assert np.abs(out["charge"] - exp["charge"]).max() <= 1e-12 * scale
        gpu.sync_size()


F32_RTOL = 1e-6  # FP32 compensated deposit vs the FP64 oracle: max |diff| / max |charge|


@pytest.mark.parametrize("nx,npart,seed", [(4, 4000, 12), (20, 1_000_000, 1)])
def test_gpu_f32_compensated_deposit_vs_fp64_oracle(nx, npart, seed):
    m = oppx.box_mesh(nx, nx, nx)
    ef = oppx.box_field_sines(m)
    pos, vel, p2c = oppx.box_particles(npart, seed, 0.01, m)
    bmin, bmax, gs = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1.0 / (2 * nx)
    ov = oppx.dh_overlay(m, bmin, bmax, gs)
    dt = 0.5 / nx / (0.01 * math.sqrt(3.0))
    gpu = oppx.FempicDevice(m, ef, pos, vel, p2c, overlay=ov, bmin=bmin, bmax=bmax, gs=gs)
    ref = oracle.FempicState(m, ef, pos, vel, p2c, overlay=ov, bmin=bmin, bmax=bmax, gs=gs)
    for s in range(3):
        gpu.step_fused(1e-3, 1.0, dt, fill="sort")
        ref.step_sorted(1e-3, 1.0, dt)
        gpu.sync_size()
        f32 = gpu.deposit_sorted_f32().cpu().numpy().astype(np.float64)
        rho64 = ref.charge
        err = np.abs(f32 - rho64).max() / np.abs(rho64).max()
        assert err <= F32_RTOL, f"step {s}: FP32 deposit error {err:.3e}"
        # deterministic: a second evaluation gives identical bits
        again = gpu.deposit_sorted_f32().cpu().numpy()
        assert np.array_equal(again, f32.astype(np.float32))
        # and the FP64 atomic deposit of the same step agrees with the oracle to 1e-12
        assert np.abs(gpu.charge.cpu().numpy() - rho64).max() <= 1e-12 * np.abs(rho64).max()


@pytest.mark.parametrize("every", [1, 2])
def test_gpu_lazy_sort_steps_equal_eager_sort_and_oracle(every):
    # Sort_Periodic with the record gather left pending and applied by the next step's move gives the
    # same state, bit for bit, as the eager sort; both follow the oracle's Sort_All steps
    nx, npart, seed = 6, 30000, 21
    m = oppx.box_mesh(nx, nx, nx)
    ef = oppx.box_field_sines(m)
    pos, vel, p2c = oppx.box_particles(npart, seed, 0.02, m)
    gs = 1.0 / 12
    ov = oppx.dh_overlay(m, (0, 0, 0), (1, 1, 1), gs)
    kw = dict(overlay=ov, bmin=(0, 0, 0), bmax=(1, 1, 1), gs=gs)
    q, ms, dt = 1e-3, 1.0, 0.5 / nx / (0.02 * math.sqrt(3.0))
    eager = oppx.FempicDevice(m, ef, pos, vel, p2c, **kw)
    lazy = oppx.FempicDevice(m, ef, pos, vel, p2c, **kw)
    orc = oracle.FempicState(m, ef, pos, vel, p2c, overlay=ov, bmin=(0, 0, 0), bmax=(1, 1, 1), gs=gs)
    for s in range(6):
        sort_now = (s + 1) % every == 0
        for g, lz in ((eager, False), (lazy, True)):
            g.step_begin()
            g.pass_fused(q, ms, dt)
            if sort_now:
                g.sort(drop_removed=True, lazy=lz)
            else:
                g.holefill()
        if sort_now:
            orc.step_sorted(q, ms, dt)
        else:
            orc.step(q, ms, dt)
        a, b, e = eager.download(), lazy.download(), orc.live()
        assert a["n"] == b["n"] == e["n"], f"step {s}"
        for key in ("p2c", "pos", "vel", "lc"):
            assert np.array_equal(a[key], b[key]), f"step {s} {key}"
            assert np.array_equal(a[key], e[key]), f"step {s} {key} vs oracle"
        assert np.abs(a["charge"] - e["charge"]).max() <= 1e-12 * np.abs(e["charge"]).max()


@pytest.mark.parametrize("nx,npart", [(3, 60000), (6, 30000)])
def test_gpu_f32_deposit_through_perm_equals_sorted(nx, npart):
    """The FP32 deposit over a lazy sort (lc read through the sort's permutation,
    oppx_fempic_deposit_sorted_f32_perm) gives the bits of the deposit after the permutation is applied:
    runs of ~190 particles per cell (nx=3, summed by whole waves) and of ~11 (nx=6, one lane each). A
    permutation entry outside the dats reads particle 0 and counts one fault."""
    m = oppx.box_mesh(nx, nx, nx)
    ef = oppx.box_field_sines(m)
    pos, vel, p2c = oppx.box_particles(npart, 13, 0.02, m)
    gs = 1.0 / 12
    ov = oppx.dh_overlay(m, (0, 0, 0), (1, 1, 1), gs)
    g = oppx.FempicDevice(m, ef, pos, vel, p2c, overlay=ov, bmin=(0, 0, 0), bmax=(1, 1, 1), gs=gs)
    g.step_begin()
    g.pass_fused(1e-3, 1.0, 0.5 / nx / (0.02 * math.sqrt(3.0)))
    g.sort(drop_removed=True, lazy=True)
    assert g.pending
    through_perm = g.deposit_sorted_f32().clone()
    g.materialize()
    sorted_lc = g.deposit_sorted_f32().clone()
    assert torch.equal(through_perm, sorted_lc)
    c = g.read_counters()
    assert c["faults"] == 0
    n = int(c["size"])
    assert n > 1000
    # a permutation with one entry past the dats (the identity otherwise: lc is now in sorted order)
    perm = torch.arange(g.cap, dtype=torch.int32, device="cuda")
    perm[n // 2] = g.cap + 5
    rc = oppx.lib().oppx_fempic_deposit_sorted_f32_perm(
        g.n_max, g.cap, g.counters.data_ptr(), g.p2c.data_ptr(), g.lc.data_ptr(), perm.data_ptr(), None, g.nc,
        g.node_start.data_ptr(), None, g.node_ref.data_ptr(), g.nn, g.cell_partials.data_ptr(), None,
        g.charge_f32.data_ptr(), g.dep_scratch.data_ptr(), g.dep_bytes, g.stream())
    assert rc == 0
    assert g.read_counters()["faults"] == 1
    # the identity permutation gives the sorted deposit again
    perm[n // 2] = n // 2
    assert oppx.lib().oppx_fempic_deposit_sorted_f32_perm(
        g.n_max, g.cap, g.counters.data_ptr(), g.p2c.data_ptr(), g.lc.data_ptr(), perm.data_ptr(), None, g.nc,
        g.node_start.data_ptr(), None, g.node_ref.data_ptr(), g.nn, g.cell_partials.data_ptr(), None,
        g.charge_f32.data_ptr(), g.dep_scratch.data_ptr(), g.dep_bytes, g.stream()) == 0
    assert torch.equal(g.charge_f32, sorted_lc)


def test_gpu_f32_deposit_reports_unsorted_p2c():
    """The FP32 deposit's precondition (p2c sorted, removed particles last) is checked: sorted input
    adds no fault; each run start that follows a larger cell or a removed particle adds one."""
    m = oppx.box_mesh(3, 3, 3)
    ef = oppx.box_field_sines(m)
    pos, vel, p2c = oppx.box_particles(3000, 5, 0.01, m)
    pos, vel, p2c = oppx.sort_particles_by_cell(pos, vel, p2c)
    gpu = oppx.FempicDevice(m, ef, pos, vel, p2c)
    gpu.deposit_sorted_f32()
    assert gpu.read_counters()["faults"] == 0
    # three descents planted in the input: a swap of two runs' neighbours, and a removed particle
    # followed by a valid cell
    q = p2c.copy()
    cuts = np.flatnonzero(np.diff(q) > 0)[[3, 20]] + 1  # first index of two runs
    for c in cuts:
        q[c - 1], q[c] = q[c], q[c - 1]  # the larger cell now comes first: one descent each
    q[100] = oppx.MAX_CELL_INDEX
    expect = int(np.sum((q[1:] < q[:-1]) & (q[1:] != oppx.MAX_CELL_INDEX)))
    g2 = oppx.FempicDevice(m, ef, pos, vel, np.where(q == oppx.MAX_CELL_INDEX, 0, q).astype(np.int32))
    g2.p2c[: len(q)] = torch.from_numpy(q.astype(np.int32)).to("cuda")
    g2.deposit_sorted_f32()
    assert expect >= 3 and g2.read_counters()["faults"] == expect
