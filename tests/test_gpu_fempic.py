"""GPU parity of the FEM-PIC hot path (C-ABI -> gfx950 kernels) against the reference fixtures
and the oracle. Bar: cell indices, hop counts, removal counts and the hole-fill permutation
bit-exact; positions, velocities and lc bit-exact (the FMAs are the reference's); FP64 atomic
deposit within 1e-12 relative to the reference's sums (atomic reordering)."""

import ctypes
import hashlib
import math

import numpy as np
import pytest

import oppx
import oracle
from conftest import golden_path

pytestmark = pytest.mark.gpu

DEPOSIT_RTOL = 1e-12  # FP64 deposit vs the reference: max |diff| / max |charge|


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load(name):
    return np.load(golden_path(name), allow_pickle=False)


def case(g):
    n = int(g["meta_nx"])
    if str(g["meta_case"]).startswith("mesh12"):  # 12-tet box read through the mesh-file loader
        nz = int(g["meta_nz"])
        m = oppx.box12_mesh(n, int(g["meta_ny"]), nz, 1.0, 1.0, nz / n)
    else:
        m = oppx.box_mesh(n, n, n)
    if "in_pos" in g:
        ef, pos, vel, p2c = g["in_c_ef"], g["in_pos"], g["in_vel"], g["in_p2c"]
    else:
        ef = oppx.box_field_sines(m)
        pos, vel, p2c = oppx.box_particles(int(g["meta_n_part"]), int(g["meta_seed"]), float(g["meta_sigma"]), m)
    kw = {}
    if int(g["meta_dh"]):
        kw = dict(overlay=g["out_overlay"], bmin=tuple(g["bmin"]), bmax=tuple(g["bmax"]), gs=float(g["meta_gs"]))
    return m, ef, pos, vel, p2c, kw


def charge_close(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= DEPOSIT_RTOL * scale


@pytest.mark.parametrize("name", ["fempic_tiny_seq.npz", "fempic_small_dh_omp1.npz", "fempic_mesh12_dh.npz"])
@pytest.mark.parametrize("mode", ["unfused", "fused", "fused-quad"])
def test_gpu_matches_reference_fixture(name, mode):
    if mode == "fused-quad":
        saved = oppx.get_tuning()
        oppx.set_tuning(**dict(saved, lanes=4))
        try:
            _fixture_run(name, "fused")
        finally:
            oppx.set_tuning(**saved)
    else:
        _fixture_run(name, mode)


def _fixture_run(name, mode):
    g = load(name)
    m, ef, pos, vel, p2c, kw = case(g)
    gpu = oppx.FempicDevice(m, ef, pos, vel, p2c, capacity=pos.shape[0] + 77, **kw)
    q, ms, dt = float(g["meta_charge"]), float(g["meta_mass"]), float(g["meta_dt"])
    steps = g["out_steps_iter_before_after_hops_maxhops"]
    for s in range(int(g["meta_steps"])):
        if mode == "unfused":
            gpu.step_unfused(q, ms, dt, record_hops=True)
        else:
            gpu.step_fused(q, ms, dt, record_hops=True)
        c = gpu.read_counters()
        assert c["faults"] == 0
        assert (c["size"], c["total_hops"], c["max_hops"]) == tuple(steps[s, 2:5]), f"step {s}: {c}"
        if s == 0:
            out = gpu.download()
            assert np.array_equal(gpu.hops[: steps[0, 1]].cpu().numpy(), g["out_s1_hops"])
            assert np.array_equal(out["p2c"], g["out_s1_p2c"])
            assert np.array_equal(out["pos"].ravel(), g["out_s1_pos"])
            assert np.array_equal(out["vel"].ravel(), g["out_s1_vel"])
            assert np.array_equal(out["lc"].ravel(), g["out_s1_lc"])
            assert charge_close(out["charge"], g["out_s1_charge"])
        gpu.sync_size()
    out = gpu.download()
    assert np.array_equal(out["p2c"], g["out_final_p2c"])
    assert np.array_equal(out["pos"].ravel(), g["out_final_pos"])
    assert np.array_equal(out["lc"].ravel(), g["out_final_lc"])
    assert charge_close(out["charge"], g["out_final_charge"])


SHAPES = [
    # cooperative tile walk (inline_hops = 0): tile sizes, refill thresholds, split relocation, persistent grid
    dict(inline_hops=0, tile=64, refill_min=64), dict(inline_hops=0, tile=128, refill_min=8),
    dict(inline_hops=0, tile=512, refill_min=32), dict(inline_hops=0, split=1),
    dict(inline_hops=0, tile=512, refill_min=1, split=1), dict(inline_hops=0, persistent=1),
    dict(inline_hops=0, tile=64, refill_min=4, split=1, persistent=1),
    # two-pass move: evaluations in the first pass before the worklist walk, block size, quad lanes
    dict(inline_hops=64), dict(inline_hops=2), dict(inline_hops=1, first_block=1024), dict(lanes=4),
    dict(lanes=4, inline_hops=1), dict(lanes=4, inline_hops=3),
]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda d: ",".join(f"{k}={v}" for k, v in d.items()))
def test_gpu_walk_launch_shapes_bit_exact(shape):
    # every launch shape of the walk gives the same cells, hops, permutation and lc
    g = load("fempic_small_dh_omp1.npz")
    m, ef, pos, vel, p2c, kw = case(g)
    saved = oppx.get_tuning()
    oppx.set_tuning(**dict(saved, **shape))
    try:
        gpu = oppx.FempicDevice(m, ef, pos, vel, p2c, **kw)
        q, ms, dt = float(g["meta_charge"]), float(g["meta_mass"]), float(g["meta_dt"])
        steps = g["out_steps_iter_before_after_hops_maxhops"]
        for s in range(int(g["meta_steps"])):
            gpu.step_fused(q, ms, dt, record_hops=True)
            c = gpu.read_counters()
            assert (c["size"], c["total_hops"], c["max_hops"], c["faults"]) == (*steps[s, 2:5], 0), f"step {s}"
            if s == 0:
                assert np.array_equal(gpu.hops[: steps[0, 1]].cpu().numpy(), g["out_s1_hops"])
            gpu.sync_size()
        out = gpu.download()
        assert np.array_equal(out["p2c"], g["out_final_p2c"]) and np.array_equal(out["lc"].ravel(), g["out_final_lc"])
        assert np.array_equal(out["pos"].ravel(), g["out_final_pos"])
        assert charge_close(out["charge"], g["out_final_charge"])
    finally:
        oppx.set_tuning(**saved)


def test_gpu_c1_scale_matches_reference():
    # 48,000 tets / 1M particles / 100 steps / 33^3 overlay, digests of the reference's final state
    g = load("fempic_c1.npz")
    m, ef, pos, vel, p2c, kw = case(g)
    gpu = oppx.FempicDevice(m, ef, pos, vel, p2c, **kw)
    q, ms, dt = float(g["meta_charge"]), float(g["meta_mass"]), float(g["meta_dt"])
    steps = g["out_steps_iter_before_after_hops_maxhops"]
    for s in range(int(g["meta_steps"])):
        gpu.step_fused(q, ms, dt, record_hops=(s == 0))
        c = gpu.read_counters()
        assert c["faults"] == 0
        assert (c["size"], c["total_hops"], c["max_hops"]) == tuple(steps[s, 2:5]), f"step {s}: {c}"
        if s == 0:
            out = gpu.download()
            assert sha(out["p2c"]) == str(g["sha_s1_p2c"])
            assert sha(gpu.hops[: steps[0, 1]].cpu().numpy()) == str(g["sha_s1_hops"])
            assert charge_close(out["charge"], g["out_s1_charge"])
        gpu.sync_size()
    out = gpu.download()
    assert sha(out["p2c"]) == str(g["sha_final_p2c"])
    assert sha(out["pos"]) == str(g["sha_final_pos"])


def test_gpu_matches_oracle_c2_shape_slice():
    # the C2 construction (100^3 cubes, 128^3 overlay) on a slice of its particles: 2 steps
    nx, sigma = 100, 0.01
    m = oppx.box_mesh(nx, nx, nx)
    ef = oppx.box_field_sines(m)
    pos, vel, p2c = oppx.box_particles(2_000_000, 2, sigma, m)
    bmin, bmax, gs = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1.0 / 128
    ov = oppx.dh_overlay(m, bmin, bmax, gs)
    dt = 0.5 / nx / (sigma * math.sqrt(3.0))
    gpu = oppx.FempicDevice(m, ef, pos, vel, p2c, overlay=ov, bmin=bmin, bmax=bmax, gs=gs)
    ref = oracle.FempicState(m, ef, pos, vel, p2c, overlay=ov, bmin=bmin, bmax=bmax, gs=gs)
    for s in range(2):
        gpu.step_fused(1e-3, 1.0, dt)
        r = ref.step(1e-3, 1.0, dt, threads=0)
        c = gpu.read_counters()
        assert c["faults"] == 0 and c["size"] == r["size"] and c["total_hops"] == r["total_hops"]
        gpu.sync_size()
    out, exp = gpu.download(), ref.live()
    assert np.array_equal(out["p2c"], exp["p2c"]) and np.array_equal(out["pos"], exp["pos"])
    assert np.array_equal(out["lc"], exp["lc"]) and np.array_equal(out["vel"], exp["vel"])
    assert charge_close(out["charge"], exp["charge"])


def test_gpu_edge_empty_and_all_removed():
    m = oppx.box_mesh(3, 3, 3)
    ef = oppx.box_field_sines(m)
    # empty set
    z = np.zeros((0, 3))
    gpu = oppx.FempicDevice(m, ef, z, z, np.zeros(0, np.int32), capacity=64)
    gpu.step_fused(1e-3, 1.0, 0.1)
    assert gpu.read_counters()["size"] == 0
    # every particle leaves the box in one step
    pos, vel, p2c = oppx.box_particles(1000, 5, 0.01, m)
    vel = np.where(pos < 0.5, -10.0, 10.0)
    gpu = oppx.FempicDevice(m, ef, pos, vel, p2c)
    gpu.step_unfused(1e-3, 1.0, 1.0)
    c = gpu.read_counters()
    assert c["size"] == 0 and c["removed"] == 1000 and c["faults"] == 0
    assert np.abs(gpu.charge.cpu().numpy()).max() == 0.0


@pytest.mark.parametrize("n,frac,seed", [(1000, 0.1, 1), (4096, 0.5, 2), (777, 0.97, 3), (70000, 0.01, 4),
                                         (64, 1.0, 5), (129, 0.0, 6)])
def test_gpu_holefill_closed_form(n, frac, seed):
    import torch

    rng = np.random.default_rng(seed)
    removed = rng.random(n) < frac
    L = oppx.lib()
    cap = n + 13
    ids = torch.zeros((1, cap), dtype=torch.float64, device="cuda")
    ids[0, :n] = torch.arange(n, dtype=torch.float64)
    p2c = torch.zeros(cap, dtype=torch.int32, device="cuda")
    p2c[:n] = torch.from_numpy(np.where(removed, oppx.MAX_CELL_INDEX, 7).astype(np.int32))
    words = int(L.oppx_hole_bits_words(n))
    bits = np.zeros(words * 64, bool)
    bits[:n] = removed
    packed = np.packbits(bits, bitorder="little").view("<i8")  # bit j of word w = particle 64w + j
    hole_bits = torch.from_numpy(packed.ravel()).to("cuda")
    ctr = torch.zeros(oppx.COUNTER_WORDS, dtype=torch.int32, device="cuda")
    ctr[0] = n
    sb = int(L.oppx_holefill_scratch_bytes(n))
    scratch = torch.zeros(sb, dtype=torch.uint8, device="cuda")
    dats = (oppx.DatDesc * 2)(oppx.DatDesc(ids.data_ptr(), 1, 8), oppx.DatDesc(p2c.data_ptr(), 1, 4))
    oppx.check(L.oppx_holefill(n, cap, ctr.data_ptr(), hole_bits.data_ptr(), dats, 2, scratch.data_ptr(), sb,
                               torch.cuda.current_stream().cuda_stream), "holefill")
    c = ctr.cpu().numpy()
    R = int(removed.sum())
    assert c[0] == n - R and c[1] == R
    marks = np.where(removed, oracle.MAX_CELL_INDEX, 7).astype(np.int32)
    from test_oracle_golden import holefill_closed_form

    expect = holefill_closed_form(marks, R)
    got = ids[0, : n - R].cpu().numpy().astype(np.int64)
    assert np.array_equal(got, expect)
    assert c[2] == int(removed[: n - R].sum())  # holes below the new size = records moved
    assert (p2c[: n - R].cpu().numpy() != oppx.MAX_CELL_INDEX).all()
