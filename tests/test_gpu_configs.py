"""Parity at the benchmark configurations themselves, not miniatures of them.

* The large single instance exactly as bench.py runs it: the 100^3-cube Morton-numbered 6 M-tet mesh,
  the 129^3 overlay, the 50 M-particle seed-2 input in cell order, and the finalize alternating hole fill
  with a Sort_Periodic sort every 2 steps whose record gather runs inside the next step's pass
  (fempic_first_kernel with GATHER), the pass before each sort leaving positions and velocities as
  48-byte records that the pass after it reads through the permutation. A 2 M-particle slice of that
  input (every 25th particle, still in cell order) runs 6 steps beside the oracle: cells, hop totals
  and maxima, positions, velocities and lc bit-exact after every step that is not left pending, FP64
  deposit within 1e-12 relative.
* The 256-instance ensemble of workloads.ensemble_specs(256) (tets 2k-60k, the particle counts scaled
  down for time): every instance's batched result equals the instance run alone, bit for bit, and 8
  instances equal the oracle, FP32 compensated deposit within 1e-6 of the oracle's FP64 charge.
* A serving stream of Poisson arrivals at 400 requests/s holding a large request (56^3 cubes =
  1,053,696 tets, 5 M particles) among the ensemble's small ones: sampled requests, the large one
  included, give the same bits alone and in the mix, and the resident bytes never exceed the budget.
"""

import math
import os
import sys

import numpy as np
import pytest

import oppx
import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

F32_RTOL = 1e-6
F64_RTOL = 1e-12
THREADS = min(16, os.cpu_count() or 1)


def charge_close(a, b, rtol=F64_RTOL):
    return np.abs(a - b).max() <= rtol * max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def bench_slice():
    import bench

    print("[configs] bench inputs: 100^3 Morton mesh, 50M particles seed 2, cell order, 129^3 overlay", flush=True)
    m, ef, pos, vel, p2c, ov, bmin, bmax, gs = bench.build_inputs(100, 50_000_000, 2, 0.01, "cell", "morton", 128)
    sl = slice(0, None, 25)
    return m, ef, pos[sl].copy(), vel[sl].copy(), p2c[sl].copy(), ov, bmin, bmax, gs


@pytest.mark.timeout(600)
@pytest.mark.parametrize("lanes", [1, 4], ids=["lane-per-particle", "quad-per-particle"])
def test_bench_pipeline_slice_matches_oracle(bench_slice, lanes):
    m, ef, pos, vel, p2c, ov, bmin, bmax, gs = bench_slice
    nx, sigma, q, ms = 100, 0.01, 1e-3, 1.0
    dt = 0.5 * (1.0 / nx) / (sigma * math.sqrt(3.0))  # bench.py's dt
    assert len(p2c) == 2_000_000 and np.all(np.diff(p2c) >= 0)
    saved = oppx.get_tuning()
    oppx.set_tuning(**dict(saved, lanes=lanes))
    try:
        _bench_pipeline(m, ef, pos, vel, p2c, ov, bmin, bmax, gs, q, ms, dt)
    finally:
        oppx.set_tuning(**saved)


def _bench_pipeline(m, ef, pos, vel, p2c, ov, bmin, bmax, gs, q, ms, dt):
    kw = dict(overlay=ov, bmin=bmin, bmax=bmax, gs=gs)
    gpu = oppx.FempicDevice(m, ef, pos, vel, p2c, **kw)
    orc = oracle.FempicState(m, ef, pos, vel, p2c, **kw)
    sort_every = 2
    for s in range(6):
        print(f"[configs] bench pipeline step {s}", flush=True)
        sorted_step = (s + 1) % sort_every == 0
        gpu.step_begin()
        # reads through the pending permutation after a sort step; before one, leaves 48-byte records
        gpu.pass_fused(q, ms, dt, stage=sorted_step)
        c = gpu.read_counters()
        if sorted_step:
            gpu.sort(drop_removed=True, lazy=True)  # applied by the next step's pass
            r = orc.step_sorted(q, ms, dt, threads=THREADS)
        else:
            gpu.holefill()
            r = orc.step(q, ms, dt, threads=THREADS)
        assert c["faults"] == 0, f"step {s}: {c}"
        assert (c["total_hops"], c["max_hops"]) == (r["total_hops"], r["max_hops"]), f"step {s}: {c} {r}"
        if sorted_step and s < 5:
            continue  # reading back would apply the pending permutation: leave it to the next pass
        out, exp = gpu.download(), orc.live()
        assert out["n"] == exp["n"], f"step {s}"
        for k in ("p2c", "pos", "vel", "lc"):
            assert np.array_equal(out[k], exp[k]), f"step {s} {k}"
        assert charge_close(out["charge"], exp["charge"]), f"step {s} charge"


def _alone_f32(d, steps):
    kw = {}
    if d.get("overlay") is not None:
        kw = dict(overlay=d["overlay"], bmin=d["bmin"], bmax=d["bmax"], gs=d["gs"])
    g = oppx.FempicDevice(d["mesh"], d["ef"], d["pos"], d["vel"], d["p2c"], **kw)
    for _ in range(steps):
        g.step_begin()
        g.pass_fused(d["charge"], d["mass"], d["dt"], deposit=False)
        g.sort(drop_removed=True)
        g.deposit_sorted_f32()
    out = g.download()
    out["charge_f32"] = g.charge_f32.cpu().numpy()
    return out


def _same(a, b):
    return a["n"] == b["n"] and all(np.array_equal(a[k], b[k]) for k in ("p2c", "pos", "vel", "lc", "charge_f32"))


@pytest.mark.timeout(900)
def test_ensemble_spec_256_batched_equals_alone_and_oracle():
    import workloads

    specs = workloads.ensemble_specs(256, particle_scale=0.02)
    tets = [6 * nx ** 3 for nx, _, _, _ in specs]
    assert len(specs) == 256 and min(tets) >= 1_500 and max(tets) <= 70_000
    print(f"[configs] ensemble: 256 instances, {sum(tets)} tets, {sum(n for _, n, _, _ in specs)} particles",
          flush=True)
    inst = [workloads.box_instance(nx, n, s, sg) for (nx, n, s, sg) in specs]
    steps = 4
    b = oppx.FempicBatch(inst)
    for _ in range(steps):
        b.step_batch(deposit="f32")
    assert b.read_counters()["faults"] == 0
    res = b.split()
    for i, d in enumerate(inst):
        a = _alone_f32(d, steps)
        assert _same(res[i], a), f"instance {i}"
    picks = np.random.default_rng(7).choice(256, 8, replace=False)
    for i in picks:
        d = inst[i]
        o = oracle.FempicState(d["mesh"], d["ef"], d["pos"], d["vel"], d["p2c"], overlay=d["overlay"], bmin=d["bmin"],
                               bmax=d["bmax"], gs=d["gs"])
        for _ in range(steps):
            o.step_sorted(d["charge"], d["mass"], d["dt"], threads=THREADS)
        e, r = o.live(), res[i]
        assert r["n"] == e["n"], f"instance {i}"
        for k in ("p2c", "pos", "vel", "lc"):
            assert np.array_equal(r[k], e[k]), f"instance {i} {k}"
        if e["n"]:
            err = np.abs(r["charge_f32"].astype(np.float64) - e["charge"]).max() / np.abs(e["charge"]).max()
            assert err <= F32_RTOL, f"instance {i}: {err}"


@pytest.mark.timeout(900)
def test_serving_stream_with_large_request_alone_equals_mix():
    import serve
    import workloads

    small = [serve.Template.of(workloads.box_instance(nx, n, s, sg))
             for (nx, n, s, sg) in workloads.ensemble_specs(24, seed=5, particle_scale=0.1)]
    large = serve.Template.of(workloads.box_instance(56, 5_000_000, 5_000_000, 0.01))
    assert large.n_cells >= 1_000_000
    rng = np.random.default_rng(5)
    reqs, t = [], 0.0
    while len(reqs) < 40:
        t += rng.exponential(1.0 / 400.0)
        tpl = large if len(reqs) == 12 else small[rng.integers(len(small))]
        reqs.append(serve.Request(len(reqs), t, tpl, 10))
    budget = 3.0 * large.device_bytes()
    arena = serve.FempicArena.for_budget(min(budget, 64e9))
    print(f"[configs] serving: {len(reqs)} requests over {t:.3f}s, budget {budget / 1e9:.2f} GB", flush=True)
    done, stats = serve.serve(arena, reqs, budget, keep_results=True)
    assert stats["requests"] == len(reqs)
    assert stats["peak_resident_bytes"] <= budget
    by_id = {r.rid: r for r in done}
    for rid in [12, 0, 5, 19, 33, 39]:
        r = by_id[rid]
        print(f"[configs] serving: request {rid} ({r.tpl.n_cells} tets, {r.tpl.n_part} particles) alone", flush=True)
        assert _same(r.result, _alone_f32(r.tpl.inst, r.steps)), f"request {rid}"
