"""Batched ensemble (one concatenated mesh + particle set, single launches over every instance):
each instance's result must equal the same instance run alone, and the oracle.

Bar: cells, positions, velocities and lc bit-exact against the alone runs and the oracle (Sort_All
finalize); FP32 compensated deposit bit-exact against the alone run (deterministic addition order)
and within 1e-6 of the oracle's FP64 charge; FP64 atomic deposit within 1e-12 relative. The captured
graph of two ensemble steps is replayed several times, with eager work and synchronisation between
replays, and every replay is checked against the eager ensemble."""

import math

import numpy as np
import pytest
import torch

import oppx
import oracle

pytestmark = pytest.mark.gpu

F32_RTOL = 1e-6
F64_RTOL = 1e-12


def make_instances(spec, seed=40):
    rng = np.random.default_rng(seed)
    out = []
    for i, (nx, npart, dh) in enumerate(spec):
        m = oppx.box_mesh(nx, nx, nx)
        ef = oppx.box_field_sines(m)
        sigma = float(rng.uniform(0.005, 0.02))
        pos, vel, p2c = oppx.box_particles(npart, 1000 + i, sigma, m)
        d = dict(mesh=m, ef=ef, pos=pos, vel=vel, p2c=p2c, charge=1e-3, mass=1.0,
                 dt=0.5 * (1.0 / nx) / (sigma * math.sqrt(3.0)))
        if dh:
            gs = 1.0 / (2 * nx)
            d.update(overlay=oppx.dh_overlay(m, (0, 0, 0), (1, 1, 1), gs), bmin=(0.0, 0.0, 0.0), bmax=(1.0, 1.0, 1.0),
                     gs=gs)
        out.append(d)
    return out


SPEC = [(3, 3000, True), (2, 500, False), (5, 12000, True), (4, 7000, True), (2, 1, True), (3, 4000, False)]


def alone(d):
    kw = {}
    if "overlay" in d:
        kw = dict(overlay=d["overlay"], bmin=d["bmin"], bmax=d["bmax"], gs=d["gs"])
    return oppx.FempicDevice(d["mesh"], d["ef"], d["pos"], d["vel"], d["p2c"], **kw)


def alone_step(g, d, deposit):
    g.step_begin()
    k_q, k_m, dt = d["charge"], d["mass"], d["dt"]
    if deposit == "f64":
        g.pass_fused(k_q, k_m, dt)
        g.sort(drop_removed=True)
    else:
        g.pass_fused(k_q, k_m, dt, deposit=False)
        g.sort(drop_removed=True)
        g.deposit_sorted_f32()


@pytest.fixture(params=[dict(), dict(inline_hops=0, persistent=1), dict(inline_hops=0, tile=64, refill_min=4, sort_items=4),
                        dict(inline_hops=1, sort_items=8), dict(lanes=4), dict(lanes=4, inline_hops=2)],
                ids=["default", "tile-persistent", "tile-small", "two-pass-1", "quad", "quad-two-pass-2"])
def tuning(request):
    saved = oppx.get_tuning()
    oppx.set_tuning(**dict(saved, **request.param))
    yield request.param
    oppx.set_tuning(**saved)


@pytest.mark.parametrize("deposit", ["f32", "f64"])
def test_batch_equals_alone_runs(tuning, deposit):
    inst = make_instances(SPEC)
    b = oppx.FempicBatch(inst)
    solo = [alone(d) for d in inst]
    for step in range(6):
        b.step_batch(deposit=deposit)
        for g, d in zip(solo, inst):
            alone_step(g, d, deposit)
        assert b.read_counters()["faults"] == 0
        parts = b.split()
        for i, (g, r) in enumerate(zip(solo, parts)):
            a = g.download()
            assert r["n"] == a["n"], f"step {step} instance {i}"
            for key in ("p2c", "pos", "vel", "lc"):
                assert np.array_equal(r[key], a[key]), f"step {step} instance {i} {key}"
            if deposit == "f32":
                assert np.array_equal(r["charge_f32"], g.charge_f32.cpu().numpy()), f"step {step} instance {i}"
            else:
                scale = max(np.abs(a["charge"]).max(), 1e-300)
                assert np.abs(r["charge"] - a["charge"]).max() <= F64_RTOL * scale


def test_batch_matches_oracle():
    inst = make_instances(SPEC, seed=41)
    b = oppx.FempicBatch(inst)
    orc = []
    for d in inst:
        kw = {}
        if "overlay" in d:
            kw = dict(overlay=d["overlay"], bmin=d["bmin"], bmax=d["bmax"], gs=d["gs"])
        orc.append(oracle.FempicState(d["mesh"], d["ef"], d["pos"], d["vel"], d["p2c"], **kw))
    for step in range(4):
        b.step_batch(deposit="f32")
        for o, d in zip(orc, inst):
            o.step_sorted(d["charge"], d["mass"], d["dt"])
        for i, (o, r) in enumerate(zip(orc, b.split())):
            e = o.live()
            assert r["n"] == e["n"]
            for key in ("p2c", "pos", "vel", "lc"):
                assert np.array_equal(r[key], e[key]), f"step {step} instance {i} {key}"
            if e["n"]:
                err = np.abs(r["charge_f32"].astype(np.float64) - e["charge"]).max() / np.abs(e["charge"]).max()
                assert err <= F32_RTOL, f"step {step} instance {i}: {err}"


def test_batch_graph_replays_match_eager():
    inst = make_instances(SPEC, seed=42)
    eager = oppx.FempicBatch(inst)
    g = oppx.FempicBatch(inst)
    g.step_batch()  # warm-up outside capture: allocations, occupancy queries
    eager.step_batch()
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            g.step_batch()
            g.step_batch()  # two steps: the sort's buffer swap returns to the starting buffers
    torch.cuda.synchronize()
    for rep in range(4):
        graph.replay()
        torch.cuda.synchronize()
        eager.step_batch()
        eager.step_batch()
        _ = torch.ones(1000, device="cuda").sum().item()  # unrelated eager work between replays
        torch.cuda.synchronize()
        a, e = g.split(), eager.split()
        for i, (x, y) in enumerate(zip(a, e)):
            assert x["n"] == y["n"], f"replay {rep} instance {i}"
            for key in ("p2c", "pos", "vel", "lc", "charge_f32"):
                assert np.array_equal(x[key], y[key]), f"replay {rep} instance {i} {key}"


@pytest.mark.parametrize("deposit", ["f32", "f64"])
def test_lazy_sequence_equals_eager_steps(tuning, deposit):
    """steps_batch (lazy sorts, 48-byte records read through the permutation, the FP32 deposit over the
    pending sort) against step_batch step by step: every step's charge, and the state after the sequence.
    The tile-walk tunings take the eager path inside steps_batch (the records need the two-pass move)."""
    inst = make_instances(SPEC, seed=43)
    lazy, eager = oppx.FempicBatch(inst), oppx.FempicBatch(inst)
    for seq, n in enumerate((3, 1, 4)):
        charges = []

        def keep(phase):
            if phase == 3:
                charges.append((lazy.charge_f32 if deposit == "f32" else lazy.charge).clone())

        print(f"[lazy] {tuning} {deposit} sequence {seq}: {n} steps", flush=True)
        lazy.steps_batch(n, deposit=deposit, mark=keep)
        assert not lazy.pending and not lazy.staged
        for s in range(n):
            eager.step_batch(deposit=deposit)
            got = charges[s].cpu().numpy()
            if deposit == "f32":
                assert np.array_equal(got, eager.charge_f32.cpu().numpy()), f"sequence {seq} step {s}"
            else:
                exp = eager.charge.cpu().numpy()
                assert np.abs(got - exp).max() <= F64_RTOL * max(np.abs(exp).max(), 1e-300), f"sequence {seq} step {s}"
        assert lazy.read_counters()["faults"] == 0
        a, e = lazy.download(), eager.download()
        assert a["n"] == e["n"], f"sequence {seq}"
        for key in ("p2c", "pos", "vel", "lc"):
            assert np.array_equal(a[key], e[key]), f"sequence {seq} {key}"


def test_lazy_sequence_graph_replays_match_eager():
    """Two lazy ensemble steps captured as one graph (they end on the starting buffers), replayed several
    times with eager work and synchronisation between replays, each replay against two eager steps."""
    inst = make_instances(SPEC, seed=44)
    eager = oppx.FempicBatch(inst)
    g = oppx.FempicBatch(inst)
    g.steps_batch(2)  # warm-up outside capture: allocations (records, swap buffers), occupancy queries
    eager.step_batch()
    eager.step_batch()
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            g.steps_batch(2)
    torch.cuda.synchronize()
    for rep in range(4):
        graph.replay()
        torch.cuda.synchronize()
        eager.step_batch()
        eager.step_batch()
        _ = torch.ones(1000, device="cuda").sum().item()  # unrelated eager work between replays
        torch.cuda.synchronize()
        a, e = g.split(), eager.split()
        for i, (x, y) in enumerate(zip(a, e)):
            assert x["n"] == y["n"], f"replay {rep} instance {i}"
            for key in ("p2c", "pos", "vel", "lc", "charge_f32"):
                assert np.array_equal(x[key], y[key]), f"replay {rep} instance {i} {key}"
