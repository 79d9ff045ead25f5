"""The C++ host API of batched stepping and continuous batching (oppx_arena_*, include/oppx.h),
driven by a C++ serving loop with no torch and no Python (tests/api/arena_app.cpp): instances of
different mesh and particle sizes arrive at different iterations, are admitted while they fit an
arena small enough that some must wait for a retirement, run their own number of steps beside the
others and are retired. Every instance's particles (positions, velocities, lc, cells, in order) and
its FP32 compensated node charge must equal the instance stepped alone (push + direct hop + walk,
Sort_All finalize, FP32 deposit) bit for bit; a few are also checked against the oracle."""

import os
import subprocess
import tempfile

import numpy as np
import pytest

import oppx
import oracle
import workloads

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "tests", "api", "arena_app")

# (nx, particles, seed, sigma, steps, arrival iteration)
SPECS = [(4, 6000, 301, 0.02, 3, 0), (6, 20000, 302, 0.01, 5, 0), (3, 2500, 303, 0.03, 2, 1),
         (7, 30000, 304, 0.015, 4, 1), (5, 12000, 305, 0.02, 6, 2), (4, 8000, 306, 0.005, 3, 4),
         (8, 40000, 307, 0.01, 2, 4), (2, 500, 308, 0.02, 4, 5)]


def _write(d, inst, steps, arrival):
    os.makedirs(d)
    m = inst["mesh"]
    dims = oppx.dh_dims(inst["bmin"], inst["bmax"], inst["gs"])
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{m['n_cells']} {m['n_nodes']} {len(inst['p2c'])} 1 {dims[0]} {dims[1]} {dims[2]} "
                f"{inst['bmin'][0]!r} {inst['bmin'][1]!r} {inst['bmin'][2]!r} {1.0 / inst['gs']!r} "
                f"{inst['charge']!r} {inst['mass']!r} {inst['dt']!r} {steps} {arrival}\n")
    arrays = dict(c_det=m["c_det"], c_vol=m["c_vol"], c2c=m["c2c"].astype(np.int32), c2n=m["c2n"].astype(np.int32),
                  c_ef=inst["ef"], pos=inst["pos"], vel=inst["vel"], p2c=inst["p2c"].astype(np.int32),
                  overlay=inst["overlay"].astype(np.int32))
    for k, a in arrays.items():
        np.ascontiguousarray(a).tofile(os.path.join(d, k + ".bin"))


def _alone(inst, steps):
    g = oppx.FempicDevice(inst["mesh"], inst["ef"], inst["pos"], inst["vel"], inst["p2c"], overlay=inst["overlay"],
                          bmin=inst["bmin"], bmax=inst["bmax"], gs=inst["gs"])
    for _ in range(steps):
        g.step_begin()
        g.pass_fused(inst["charge"], inst["mass"], inst["dt"], deposit=False)
        g.sort(drop_removed=True)
        g.deposit_sorted_f32()
    out = g.download()
    out["charge_f32"] = g.charge_f32.cpu().numpy()
    return out


@pytest.mark.timeout(300)
def test_cpp_arena_serving_loop_equals_alone():
    if not os.path.exists(APP):
        pytest.skip("tests/api/arena_app is not built (make -C tests/api)")
    inst = [workloads.box_instance(nx, n, seed, sigma) for nx, n, seed, sigma, _, _ in SPECS]
    cells = sum(int(d["mesh"]["n_cells"]) for d in inst)
    parts = sum(len(d["p2c"]) for d in inst)
    with tempfile.TemporaryDirectory() as td:
        for i, (d, spec) in enumerate(zip(inst, SPECS)):
            _write(os.path.join(td, f"inst{i}"), d, spec[4], spec[5])
        # about half the total: later arrivals wait for retirements; at most 4 resident at once
        cap_cells, cap_parts = max(cells // 2, 6 * 8 ** 3), max(parts // 2, 40000)
        args = [APP, td, str(len(inst)), str(cap_cells), str(cap_cells), str(cap_parts), str(cap_cells * 2), "4"]
        print(f"[arena] {' '.join(args[2:])}", flush=True)
        r = subprocess.run(args, capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, f"rc={r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
        waited = 0
        for i, (d, spec) in enumerate(zip(inst, SPECS)):
            o = os.path.join(td, f"out{i}")
            k, admitted, retired = (int(v) for v in open(os.path.join(o, "n.txt")).read().split())
            assert retired - admitted == spec[4] - 1, f"instance {i} ran {retired - admitted + 1} steps"
            waited += admitted > spec[5]
            got = dict(pos=np.fromfile(os.path.join(o, "pos.bin")).reshape(-1, 3),
                       vel=np.fromfile(os.path.join(o, "vel.bin")).reshape(-1, 3),
                       lc=np.fromfile(os.path.join(o, "lc.bin")).reshape(-1, 4),
                       p2c=np.fromfile(os.path.join(o, "p2c.bin"), dtype=np.int32),
                       charge_f32=np.fromfile(os.path.join(o, "charge_f32.bin"), dtype=np.float32))
            exp = _alone(d, spec[4])
            assert k == exp["n"], f"instance {i}: {k} particles, alone {exp['n']}"
            for key in ("pos", "vel", "lc", "p2c", "charge_f32"):
                assert np.array_equal(got[key], exp[key]), f"instance {i} {key}"
            if i in (1, 3, 7):
                orc = oracle.FempicState(d["mesh"], d["ef"], d["pos"], d["vel"], d["p2c"], overlay=d["overlay"],
                                         bmin=d["bmin"], bmax=d["bmax"], gs=d["gs"])
                for _ in range(spec[4]):
                    orc.step_sorted(d["charge"], d["mass"], d["dt"])
                e = orc.live()
                for key in ("pos", "vel", "lc", "p2c"):
                    assert np.array_equal(got[key], e[key]), f"instance {i} {key} vs oracle"
        assert waited >= 1, "the arena was large enough for every arrival: no admission waited"


def _desc(d, keep):
    """oppx_instance of a workloads.box_instance dict (the numpy arrays kept alive in `keep`)."""
    m = d["mesh"]
    arrs = [np.ascontiguousarray(a) for a in (m["c_det"], m["c_vol"], m["c2c"].astype(np.int32),
                                               m["c2n"].astype(np.int32), d["ef"], d["pos"], d["vel"],
                                               d["p2c"].astype(np.int32), d["overlay"].astype(np.int32))]
    keep.extend(arrs)
    p = [a.ctypes.data for a in arrs]
    dims = oppx.dh_dims(d["bmin"], d["bmax"], d["gs"])
    import ctypes
    return oppx.Instance(int(m["n_cells"]), int(m["n_nodes"]), len(d["p2c"]), *p, (ctypes.c_int64 * 3)(*dims),
                         (ctypes.c_double * 3)(*d["bmin"]), 1.0 / d["gs"], d["charge"], d["mass"], d["dt"])


def test_arena_error_paths_leave_it_usable():
    """Out-of-range indices are refused before anything is copied; a retirement before the step that
    follows an admission is refused (the set is not in cell order yet); an instance that does not fit is
    refused with OPPX_ERR_NOMEM until a retirement frees room; a retired slot cannot be retired again; an
    instance with no particles is served. After each refusal the arena still serves correctly."""
    import ctypes

    L = oppx.lib()
    keep = []
    a_inst = workloads.box_instance(4, 5000, 401, 0.02)
    b_inst = workloads.box_instance(5, 9000, 402, 0.01)
    empty = workloads.box_instance(3, 1, 403, 0.02)
    empty = dict(empty, pos=empty["pos"][:0], vel=empty["vel"][:0], p2c=empty["p2c"][:0])
    nc = {k: int(d["mesh"]["n_cells"]) for k, d in (("a", a_inst), ("b", b_inst), ("e", empty))}
    cells = nc["e"] + nc["b"]  # b fits only where a was, once a has gone
    arena = ctypes.c_void_p()
    assert L.oppx_arena_create(cells, cells, 20000, 20000, 4, None, ctypes.byref(arena)) == oppx.OK

    def retire(slot, d):
        cap = max(1, len(d["p2c"]))
        n = ctypes.c_int64()
        out = {k: np.zeros((cap, w)) for k, w in (("pos", 3), ("vel", 3), ("lc", 4))}
        p2c, q = np.zeros(cap, np.int32), np.zeros(int(d["mesh"]["n_nodes"]), np.float32)
        rc = L.oppx_arena_retire(arena, slot, cap, ctypes.byref(n), out["pos"].ctypes.data, out["vel"].ctypes.data,
                                 out["lc"].ctypes.data, p2c.ctypes.data, q.ctypes.data)
        k = n.value
        return rc, dict(n=k, pos=out["pos"][:k], vel=out["vel"][:k], lc=out["lc"][:k], p2c=p2c[:k], charge_f32=q)

    def admit(d):
        slot = ctypes.c_int32(-1)
        return L.oppx_arena_admit(arena, ctypes.byref(_desc(d, keep)), ctypes.byref(slot)), slot.value

    try:
        bad = dict(a_inst, p2c=np.where(np.arange(len(a_inst["p2c"])) == 7, 10 ** 6, a_inst["p2c"]).astype(np.int32))
        assert admit(bad)[0] == oppx.ERR_ARG
        rc, se = admit(empty)
        assert rc == oppx.OK
        rc, sa = admit(a_inst)
        assert rc == oppx.OK
        assert admit(b_inst)[0] == oppx.ERR_NOMEM  # no room beside a: refused, nothing changed
        assert retire(sa, a_inst)[0] == oppx.ERR_STATE  # admitted since the last step
        for _ in range(3):
            assert L.oppx_arena_step(arena) == oppx.OK
        rc, got = retire(sa, a_inst)
        assert rc == oppx.OK
        assert retire(sa, a_inst)[0] == oppx.ERR_ARG  # already retired
        exp = _alone(a_inst, 3)
        assert got["n"] == exp["n"]
        for key in ("pos", "vel", "lc", "p2c", "charge_f32"):
            assert np.array_equal(got[key], exp[key]), key
        rc, sb = admit(b_inst)  # a's cells are free again
        assert rc == oppx.OK
        for _ in range(2):
            assert L.oppx_arena_step(arena) == oppx.OK
        rc, got = retire(sb, b_inst)
        assert rc == oppx.OK
        exp = _alone(b_inst, 2)
        assert got["n"] == exp["n"]
        for key in ("pos", "vel", "lc", "p2c", "charge_f32"):
            assert np.array_equal(got[key], exp[key]), f"b {key}"
        rc, got = retire(se, empty)
        assert rc == oppx.OK and got["n"] == 0 and not got["charge_f32"].any()
        c = oppx.MoveCounters()
        assert L.oppx_arena_counters(arena, ctypes.byref(c)) == oppx.OK and c.faults == 0
    finally:
        assert L.oppx_arena_destroy(arena) == oppx.OK
