"""The fused step with positions and velocities staged as 48-byte records (oppx_fempic_step_fused_rec,
oppx_gather_records) against the oracle, in every order a caller can reach them in: a staged pass
followed by a lazy sort and a pass that reads the records through the permutation (the bench's
Sort_Periodic pipeline), staged passes back to back with a sort after each (records read and written
by the same pass, two buffers in turn), a staged pass read back or hole-filled without a sort, a staged
pass followed by an eager sort, and a lazy sort of staged records read back before the next pass. The
worklist walk reads the staged position when the first pass hands a particle on (inline_hops 1), and
the quad-lane first pass stages the same records. Cells, hop totals, positions, velocities and lc
bit-exact; FP64 deposit within 1e-12 relative."""

import math

import numpy as np
import pytest

import oppx
import oracle

pytestmark = pytest.mark.gpu

# what follows each staged pass: (lazy sort | eager sort | hole fill), and whether the state is read back
SCHEDULES = {
    "periodic-lazy": ["stage-lazy", "plain-holefill"] * 3,
    "every-step-lazy": ["stage-lazy"] * 4 + ["plain-holefill"],
    "stage-then-holefill": ["stage-holefill", "stage-holefill", "plain-holefill"],
    "stage-then-eager-sort": ["stage-eager", "plain-holefill", "stage-eager"],
    "lazy-read-back": ["stage-lazy-read", "stage-lazy", "plain-holefill"],
}


@pytest.fixture(params=[dict(lanes=1), dict(lanes=4), dict(lanes=1, inline_hops=1)],
                ids=["lane-per-particle", "quad-per-particle", "worklist-after-one-hop"])
def shape(request):
    saved = oppx.get_tuning()
    oppx.set_tuning(**dict(saved, **request.param))
    yield request.param
    oppx.set_tuning(**saved)


@pytest.mark.parametrize("schedule", sorted(SCHEDULES))
def test_staged_records_match_oracle(schedule, shape):
    nx, n, sigma, seed = 6, 30000, 0.02, 211
    m = oppx.box_mesh(nx, nx, nx, order="morton")
    ef = oppx.box_field_sines(m)
    pos, vel, p2c = oppx.box_particles(n, seed, sigma, m)
    pos, vel, p2c = oppx.sort_particles_by_cell(pos, vel, p2c)
    gs = 1.0 / (2 * nx)
    kw = dict(overlay=oppx.dh_overlay(m, (0, 0, 0), (1, 1, 1), gs), bmin=(0, 0, 0), bmax=(1, 1, 1), gs=gs)
    gpu = oppx.FempicDevice(m, ef, pos, vel, p2c, **kw)
    orc = oracle.FempicState(m, ef, pos, vel, p2c, **kw)
    q, ms, dt = 1e-3, 1.0, 0.5 / nx / (sigma * math.sqrt(3.0))
    steps = SCHEDULES[schedule]
    for s, what in enumerate(steps):
        print(f"[staged] {schedule} {shape} step {s}: {what}", flush=True)
        gpu.step_begin()
        gpu.pass_fused(q, ms, dt, stage=what.startswith("stage"))
        assert gpu.staged == what.startswith("stage"), f"step {s}"
        c = gpu.read_counters()
        if "holefill" in what:
            gpu.holefill()
            r = orc.step(q, ms, dt)
        else:
            gpu.sort(drop_removed=True, lazy="lazy" in what)
            r = orc.step_sorted(q, ms, dt)
        assert c["faults"] == 0 and c["total_hops"] == r["total_hops"] and c["max_hops"] == r["max_hops"], f"step {s}"
        last = s == len(steps) - 1
        if "lazy" in what and not last and not what.endswith("read"):
            continue  # left to the next pass, which reads the records through the permutation
        out, exp = gpu.download(), orc.live()
        assert not gpu.staged and not gpu.pending
        assert out["n"] == exp["n"], f"step {s}"
        for k in ("p2c", "pos", "vel", "lc"):
            assert np.array_equal(out[k], exp[k]), f"step {s} {k}"
        scale = max(np.abs(exp["charge"]).max(), 1e-300)
        assert np.abs(out["charge"] - exp["charge"]).max() <= 1e-12 * scale, f"step {s} charge"
