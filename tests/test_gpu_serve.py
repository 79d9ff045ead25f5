"""Serving loop (continuous batching) on the GPU: every request's result is the same whether it ran
alone or in the mix, and admission keeps the resident state within the HBM budget.

Bar: a request's particles (cells, positions, velocities, lc) bit-exact against the same instance
stepped alone with the same step sequence (batched move, Sort_All finalize, FP32 compensated deposit),
FP32 node charge bit-exact (the addition order is fixed by the data), and the alone run against the
oracle within 1e-6 relative (FP32 compensated vs FP64)."""

import math

import numpy as np
import pytest

import oppx
import oracle
import serve
import workloads

pytestmark = pytest.mark.gpu


def make_templates(specs, seed=60):
    rng = np.random.default_rng(seed)
    out = []
    for i, (nx, n) in enumerate(specs):
        d = workloads.box_instance(nx, n, seed * 1000 + i, float(rng.uniform(0.005, 0.02)))
        out.append(serve.Template.of(d))
    return out


def alone(tpl, steps):
    d = tpl.inst
    g = oppx.FempicDevice(d["mesh"], d["ef"], d["pos"], d["vel"], d["p2c"], overlay=d["overlay"], bmin=d["bmin"],
                          bmax=d["bmax"], gs=d["gs"])
    for _ in range(steps):
        g.step_begin()
        g.pass_fused(d["charge"], d["mass"], d["dt"], deposit=False)
        g.sort(drop_removed=True)
        g.deposit_sorted_f32()
    out = g.download()
    out["charge_f32"] = g.charge_f32.cpu().numpy()
    return out


def same(a, b):
    return a["n"] == b["n"] and all(np.array_equal(a[k], b[k]) for k in ("p2c", "pos", "vel", "lc", "charge_f32"))


def test_requests_alone_and_in_the_mix_are_identical():
    tpls = make_templates([(3, 4000), (5, 20000), (2, 700), (4, 9000), (6, 30000), (3, 1)])
    steps = [4, 6, 3, 5, 4, 2]
    # staggered arrivals: later requests join while earlier ones are mid-way
    reqs = [serve.Request(i, 0.0 if i < 2 else 1e-3 * i, t, s) for i, (t, s) in enumerate(zip(tpls, steps))]
    arena = serve.FempicArena(max_cells=20000, max_nodes=5000, max_particles=120000, max_overlay=40000, max_insts=8)
    done, stats = serve.serve(arena, reqs, budget_bytes=1e12, keep_results=True)
    assert stats["requests"] == len(reqs)
    for r in done:
        ref = alone(r.tpl, r.steps)
        assert same(r.result, ref), f"request {r.rid}"


def test_alone_run_matches_oracle():
    (tpl,) = make_templates([(4, 15000)], seed=61)
    steps = 5
    out = alone(tpl, steps)
    d = tpl.inst
    o = oracle.FempicState(d["mesh"], d["ef"], d["pos"], d["vel"], d["p2c"], overlay=d["overlay"], bmin=d["bmin"],
                           bmax=d["bmax"], gs=d["gs"])
    for _ in range(steps):
        o.step_sorted(d["charge"], d["mass"], d["dt"])
    e = o.live()
    for k in ("p2c", "pos", "vel", "lc"):
        assert np.array_equal(out[k], e[k]), k
    err = np.abs(out["charge_f32"].astype(np.float64) - e["charge"]).max() / np.abs(e["charge"]).max()
    assert err <= 1e-6


def test_admission_waits_for_budget():
    tpls = make_templates([(4, 20000)] * 4, seed=62)
    need = tpls[0].device_bytes()
    reqs = [serve.Request(i, 0.0, t, 3) for i, t in enumerate(tpls)]
    arena = serve.FempicArena(max_cells=30000, max_nodes=8000, max_particles=200000, max_overlay=40000, max_insts=8)
    done, stats = serve.serve(arena, reqs, budget_bytes=2.5 * need, keep_results=True)
    assert stats["requests"] == 4 and stats["peak_resident_bytes"] <= 2.5 * need
    # two fit at once: the third is admitted only after the first pair retires (frees its ranges and
    # budget; its results reach the host while the next step runs)
    by_id = {r.rid: r for r in done}
    assert by_id[2].t_admit >= min(by_id[0].t_retire, by_id[1].t_retire)
    assert all(r.t_done >= r.t_retire for r in done)
    for r in done:
        assert same(r.result, alone(r.tpl, r.steps)), f"request {r.rid}"


def test_resident_templates_and_plain_clones_give_the_same_results():
    """Templates kept in HBM (admission copies device to device) and the plain path's prototypes cloned on
    the device for every request: the same bits as the host-staged admission and a fresh alone run."""
    tpls = make_templates([(3, 4000), (5, 20000), (4, 9000), (2, 700)], seed=63)
    for t in tpls:
        t.make_resident()
        t.prototype()
    steps = [3, 5, 4, 2]
    reqs = [serve.Request(i, 0.0 if i < 2 else 1e-3 * i, t, s) for i, (t, s) in enumerate(zip(tpls + tpls, steps * 2))]
    arena = serve.FempicArena(max_cells=20000, max_nodes=5000, max_particles=120000, max_overlay=40000, max_insts=8)
    done, stats = serve.serve(arena, reqs, budget_bytes=1e12, keep_results=True)
    assert stats["requests"] == len(reqs)
    plain = [serve.Request(r.rid, 0.0, r.tpl, r.steps) for r in reqs]
    serve.serve_plain(plain, keep_results=True)
    for r, p in zip(sorted(done, key=lambda r: r.rid), plain):
        ref = alone(r.tpl, r.steps)
        assert same(r.result, ref), f"request {r.rid} (batched, resident template)"
        assert same(p.result, ref), f"request {r.rid} (plain, cloned prototype)"
    # the prototypes were never stepped: a clone still starts from the template's initial state
    for t in tpls:
        g = t.prototype().clone()
        out = g.download()
        assert out["n"] == t.n_part and np.array_equal(out["pos"], t.inst["pos"])
