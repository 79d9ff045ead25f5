"""Serving loop for independent FEM-PIC simulation segments on one GPU: continuous batching.

Requests (an instance plus a number of steps) are admitted into a resident *arena* as soon as they
have arrived and their state fits both the arena's free ranges and the HBM budget; every server
iteration then runs ONE batched step over all resident requests (batched push/move, Sort_All
finalize, FP32 compensated deposit: a fixed number of launches whatever the mix), and requests that
have done their steps are handed back and their ranges freed at once, so later arrivals join the
running batch between iterations instead of waiting for it to drain. The bulk of a retired request's
state goes back to the host on a copy stream while the next iteration's step runs.

Arena layout (device): cells, nodes and overlay words are ranges allocated first-fit per request;
every cell, node and p2c index is stored concatenated (offset by the request's base); particles of
all requests share one particle set that the per-step sort keeps ordered by concatenated cell, so a
request's particles are one contiguous range found by a lower-bound search. A retired request's
particles are marked removed and leave the set at the next sort. Per-request results are identical
to the request run alone (tests/test_gpu_serve.py).
"""

from __future__ import annotations

import ctypes
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

import oppx
from oppx import MAX_CELL_INDEX, BatchInst, DatDesc, FempicParts, TetMesh, check, lib


class RangeAlloc:
    """First-fit allocator of index ranges in [0, cap)."""

    def __init__(self, cap: int):
        self.cap = int(cap)
        self.free = [(0, self.cap)]  # sorted, disjoint (start, length)
        self.used = {}               # start -> length

    def alloc(self, n: int):
        n = int(n)
        if n <= 0:
            return None
        for k, (s, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (s + n, ln - n)
                self.used[s] = n
                return s
        return None

    def release(self, start: int):
        n = self.used.pop(start)
        self.free.append((start, n))
        self.free.sort()
        merged = []
        for s, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((s, ln))
        self.free = merged

    def extent(self) -> int:
        """End of the highest allocated range (kernels cover [0, extent))."""
        return max((s + n for s, n in self.used.items()), default=0)

    def in_use(self) -> int:
        return sum(self.used.values())


def _aux_bytes(vol: np.ndarray, c2c_global: np.ndarray) -> np.ndarray:
    """oppx_tet_aux records {c2 = (1/6)/volume, c2c[4], pad[2]} (32 B each) as float64 words."""
    n = len(vol)
    rec = np.zeros(n, dtype=[("c2", "<f8"), ("c2c", "<i4", (4,)), ("pad", "<i4", (2,))])
    rec["c2"] = (1.0 / 6.0) / vol  # the same IEEE division oppx_tet_aux_build performs
    rec["c2c"] = c2c_global
    return rec.view(np.float64).reshape(-1)


@dataclass
class Template:
    """One instance, staged once in pinned host memory in the arena's layout (indices local to the
    instance) so an admission is a handful of asynchronous copies plus index rebasing on the device."""
    inst: dict
    n_cells: int
    n_nodes: int
    n_part: int
    overlay_words: int
    host: dict = field(default_factory=dict)
    dev: dict = field(default_factory=dict)  # the same arrays resident in HBM (make_resident)
    proto: object = None                     # a FempicDevice of the instance, never stepped (prototype)

    @staticmethod
    def of(inst: dict) -> "Template":
        import torch
        m = inst["mesh"]
        nc, nn = int(m["n_cells"]), int(m["n_nodes"])
        if not ((m["c2c"] >= -1) & (m["c2c"] < nc)).all() or not ((m["c2n"] >= 0) & (m["c2n"] < nn)).all():
            raise ValueError("mesh maps out of range")
        if len(inst["p2c"]) and not ((inst["p2c"] >= 0) & (inst["p2c"] < nc)).all():
            raise ValueError("p2c out of range")
        ov = inst.get("overlay")
        if ov is not None and not (((ov >= 0) & (ov < nc)) | (ov == MAX_CELL_INDEX)).all():
            raise ValueError("overlay entries out of range")
        st, ref = oppx.node_cell_csr(m)
        pin = lambda a: torch.from_numpy(np.ascontiguousarray(a)).pin_memory()
        host = dict(det=pin(m["c_det"].reshape(-1)), aux=pin(_aux_bytes(m["c_vol"], m["c2c"].astype(np.int32))),
                    c2n=pin(m["c2n"].reshape(-1).astype(np.int32)), ef=pin(inst["ef"].reshape(-1)),
                    nstart=pin(st[:-1].astype(np.int32)), nend=pin(st[1:].astype(np.int32)),
                    nref=pin(ref.astype(np.int32)), pos=pin(inst["pos"].T), vel=pin(inst["vel"].T),
                    p2c=pin(inst["p2c"].astype(np.int32)))
        if ov is not None:
            host["overlay"] = pin(ov.astype(np.int32))
        return Template(inst, nc, nn, len(inst["p2c"]), 0 if ov is None else int(ov.size), host)

    def make_resident(self, device="cuda"):
        """Keep the staged arrays in HBM as well: an admission then copies device to device instead of
        over the host link (the template pool stays resident across requests)."""
        if not self.dev:
            self.dev = {k: v.to(device, non_blocking=True) for k, v in self.host.items()}
        return self

    def resident_bytes(self) -> int:
        return sum(v.numel() * v.element_size() for v in self.dev.values())

    def prototype(self):
        """The instance as its own FempicDevice (built once, never stepped): the plain per-request path
        clones it (device copies of the particle state, the mesh shared read-only) for every request."""
        if self.proto is None:
            d = self.inst
            self.proto = oppx.FempicDevice(d["mesh"], d["ef"], d["pos"], d["vel"], d["p2c"], overlay=d.get("overlay"),
                                           bmin=d.get("bmin"), bmax=d.get("bmax"), gs=d.get("gs"))
        return self.proto

    def device_bytes(self) -> int:
        """HBM a resident request holds: mesh records, node rows, overlay, particles (+ sort buffers)."""
        per_cell = 128 + 32 + 16 + 24 + 4 + 16 + 16  # det, aux, c2n, ef, cell_inst, node_ref, partials
        per_node = 8 + 4 + 8 + 4                     # rows, charges
        per_part = 2 * 84 + 8 + 8 + 4                # state + swap, sort keys, worklist, hops
        return per_cell * self.n_cells + per_node * self.n_nodes + 4 * self.overlay_words + per_part * self.n_part


class FempicArena:
    """Resident state of the serving loop (see the module docstring)."""

    def __init__(self, max_cells: int, max_nodes: int, max_particles: int, max_overlay: int, max_insts: int = 4096,
                 device="cuda"):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        L = lib()
        check(L.oppx_init(self.dev.index or 0), "oppx_init")
        z = lambda n, dt: torch.zeros(int(n), dtype=dt, device=self.dev)
        C, N, P = int(max_cells), int(max_nodes), int(max_particles)
        self.det, self.aux, self.c2n, self.ef = z(16 * C, torch.float64), z(4 * C, torch.float64), \
            z(4 * C, torch.int32), z(3 * C, torch.float64)
        self.cell_inst = torch.full((C,), -1, dtype=torch.int32, device=self.dev)
        self.overlay = z(max(1, max_overlay), torch.int32)
        self.node_start, self.node_end, self.node_ref = z(N, torch.int32), z(N, torch.int32), z(4 * C, torch.int32)
        self.charge_f32, self.cell_partials = z(N, torch.float32), z(4 * C, torch.float32)
        self.inst_table = z(max_insts * ctypes.sizeof(BatchInst), torch.uint8)
        # pinned host mirrors: records and the live count are copied asynchronously from here (a slot's
        # record is rewritten only after its request retired, which synchronises)
        self.h_table = torch.zeros(max_insts * ctypes.sizeof(BatchInst), dtype=torch.uint8, pin_memory=True)
        self.h_size = torch.zeros(1, dtype=torch.int32, pin_memory=True)
        self.pos, self.vel, self.lc = (torch.zeros((d, P), dtype=torch.float64, device=self.dev) for d in (3, 3, 4))
        self.p2c = torch.full((P,), MAX_CELL_INDEX, dtype=torch.int32, device=self.dev)
        self.swap = [torch.empty_like(t) for t in (self.pos, self.vel, self.lc, self.p2c)]
        self.counters = z(oppx.COUNTER_WORDS, torch.int32)
        self.hole_bits = z(int(L.oppx_hole_bits_words(P)), torch.int64)
        self.work_bytes = int(L.oppx_fempic_work_bytes(P))
        self.work = z(max(self.work_bytes, 8), torch.uint8)
        self.sort_bytes = int(L.oppx_sort_scratch_bytes(P))
        self.sort_scratch = z(self.sort_bytes, torch.uint8)
        self.dep_bytes = int(L.oppx_fempic_deposit_scratch_bytes(P))
        self.dep_scratch = z(self.dep_bytes, torch.uint8)
        self.cells, self.nodes, self.ovl = RangeAlloc(C), RangeAlloc(N), RangeAlloc(max(1, max_overlay))
        self.slots_free = list(range(max_insts - 1, -1, -1))
        self.res = {}  # slot -> (template, cb, nb, ob)
        self.dropping = 0  # particles of retired requests still in the set (marked removed) until the next sort
        # results of retired requests go to the host on their own stream, beside the next step
        self.copy_stream = torch.cuda.Stream(device=self.dev)
        self.cap, self.size = P, 0

    def stream(self):
        return self.torch.cuda.current_stream(self.dev).cuda_stream

    def _t(self, a, dt):
        return self.torch.from_numpy(np.ascontiguousarray(a)).to(self.dev, dtype=dt)

    def can_admit(self, tpl: Template) -> bool:
        return (bool(self.slots_free) and self.size + tpl.n_part <= self.cap and
                any(ln >= tpl.n_cells for _, ln in self.cells.free) and
                any(ln >= tpl.n_nodes for _, ln in self.nodes.free) and
                (tpl.overlay_words == 0 or any(ln >= tpl.overlay_words for _, ln in self.ovl.free)))

    def admit(self, tpl: Template):
        """Copy a request's staged mesh and particles into free ranges (asynchronously, on the
        arena's stream) and rebase its indices on the device; returns its slot or None."""
        if not self.can_admit(tpl):
            return None
        L, s = lib(), self.stream()
        torch = self.torch
        h, d = (tpl.dev or tpl.host), tpl.inst  # device to device when the template is resident
        nc, nn, n = tpl.n_cells, tpl.n_nodes, tpl.n_part
        cb, nb = self.cells.alloc(nc), self.nodes.alloc(nn)
        ob = self.ovl.alloc(tpl.overlay_words) if tpl.overlay_words else -1
        slot = self.slots_free.pop()
        cp = lambda dst, src: dst.copy_(src, non_blocking=True)
        cp(self.det[16 * cb:16 * (cb + nc)], h["det"])
        cp(self.aux[4 * cb:4 * (cb + nc)], h["aux"])
        cp(self.c2n[4 * cb:4 * (cb + nc)], h["c2n"])
        cp(self.ef[3 * cb:3 * (cb + nc)], h["ef"])
        cp(self.node_start[nb:nb + nn], h["nstart"])
        cp(self.node_end[nb:nb + nn], h["nend"])
        cp(self.node_ref[4 * cb:4 * (cb + nc)], h["nref"])
        i4 = lambda t, off: t.data_ptr() + 4 * off
        check(L.oppx_offset_indices(i4(self.aux, 8 * cb), nc, 8, 2, 4, cb, s), "rebase c2c")  # aux: c2c at ints 2..5
        check(L.oppx_offset_indices(i4(self.c2n, 4 * cb), 4 * nc, 1, 0, 1, nb, s), "rebase c2n")
        check(L.oppx_offset_indices(i4(self.node_start, nb), nn, 1, 0, 1, 4 * cb, s), "rebase rows")
        check(L.oppx_offset_indices(i4(self.node_end, nb), nn, 1, 0, 1, 4 * cb, s), "rebase rows")
        check(L.oppx_offset_indices(i4(self.node_ref, 4 * cb), 4 * nc, 1, 0, 1, 4 * cb, s), "rebase refs")
        pat = np.array([slot], np.int32)
        check(L.oppx_fill_pattern(i4(self.cell_inst, cb), nc, 1, 4, pat.ctypes.data, 0, s), "cell_inst")
        rec = BatchInst()
        rec.cell_base, rec.n_cells, rec.node_base, rec.n_nodes = cb, nc, nb, nn
        if ob >= 0:
            cp(self.overlay[ob:ob + tpl.overlay_words], h["overlay"])
            check(L.oppx_offset_indices(i4(self.overlay, ob), tpl.overlay_words, 1, 0, 1, cb, s), "rebase overlay")
            dims = oppx.dh_dims(d["bmin"], d["bmax"], d["gs"])
            rec.dh_cell = self.overlay.data_ptr() + 4 * ob
            for a in range(3):
                rec.dims[a], rec.min[a] = int(dims[a]), float(d["bmin"][a])
            rec.inv_gs = 1.0 / float(d["gs"])
        rec.dt = float(d["dt"])
        rec.k = float(d["charge"]) / float(d["mass"]) * float(d["dt"])
        sz = ctypes.sizeof(BatchInst)
        self.h_table[slot * sz:(slot + 1) * sz] = torch.frombuffer(bytearray(bytes(rec)), dtype=torch.uint8)
        self.inst_table[slot * sz:(slot + 1) * sz].copy_(self.h_table[slot * sz:(slot + 1) * sz], non_blocking=True)
        n0 = self.size
        if n:
            cp(self.pos[:, n0:n0 + n], h["pos"])
            cp(self.vel[:, n0:n0 + n], h["vel"])
            cp(self.p2c[n0:n0 + n], h["p2c"])
            check(L.oppx_offset_indices(i4(self.p2c, n0), n, 1, 0, 1, cb, s), "rebase p2c")
            zero = np.zeros(1, np.float64)
            check(L.oppx_fill_pattern(self.lc.data_ptr() + 8 * n0, n, 4, 8, zero.ctypes.data, self.cap, s), "lc")
        self.size = n0 + n
        self.res[slot] = (tpl, cb, nb, ob)
        return slot

    def publish_size(self):
        """Make the device live count match the host's after admissions (one small asynchronous copy;
        the pinned value is rewritten only after the step's closing synchronisation)."""
        self.h_size[0] = self.size
        self.counters[0:1].copy_(self.h_size, non_blocking=True)

    def step(self):
        """One batched step of every resident request: push/move, Sort_All finalize, FP32 deposit."""
        L, s = lib(), self.stream()
        n_cells, n_nodes = max(1, self.cells.extent()), max(1, self.nodes.extent())
        mesh = TetMesh(n_cells, n_nodes, self.det.data_ptr(), self.aux.data_ptr(), self.c2n.data_ptr(),
                       self.ef.data_ptr())
        parts = FempicParts(self.cap, self.pos.data_ptr(), self.vel.data_ptr(), self.lc.data_ptr(),
                            self.p2c.data_ptr())
        n_insts = max(self.res) + 1 if self.res else 1
        self.publish_size()
        check(L.oppx_fempic_step_begin(self.counters.data_ptr(), None, 0, s), "step_begin")
        if self.size:
            check(L.oppx_fempic_batch_step_fused(self.size, ctypes.byref(parts), ctypes.byref(mesh),
                                                 self.cell_inst.data_ptr(), self.inst_table.data_ptr(), n_insts,
                                                 self.counters.data_ptr(), self.hole_bits.data_ptr(),
                                                 self.work.data_ptr(), self.work_bytes, None, None, s),
                  "batch_step")
        cur = [self.pos, self.vel, self.lc, self.p2c]
        dims = [(3, 8), (3, 8), (4, 8), (1, 4)]
        ins = (DatDesc * 4)(*[DatDesc(t.data_ptr(), dd, e) for t, (dd, e) in zip(cur, dims)])
        outs = (DatDesc * 4)(*[DatDesc(t.data_ptr(), dd, e) for t, (dd, e) in zip(self.swap, dims)])
        check(L.oppx_sort_by_cell(self.size, self.cap, self.counters.data_ptr(), n_cells, self.p2c.data_ptr(), ins,
                                  outs, 4, None, 1, self.sort_scratch.data_ptr(), self.sort_bytes, s), "sort")
        self.pos, self.vel, self.lc, self.p2c = self.swap
        self.swap = cur
        check(L.oppx_fempic_deposit_sorted_f32(self.size, self.cap, self.counters.data_ptr(), self.p2c.data_ptr(),
                                               self.lc.data_ptr(), n_cells, self.node_start.data_ptr(),
                                               self.node_end.data_ptr(), self.node_ref.data_ptr(), n_nodes,
                                               self.cell_partials.data_ptr(), None, self.charge_f32.data_ptr(),
                                               self.dep_scratch.data_ptr(), self.dep_bytes, s),
              "deposit_f32")
        self.size = int(self.counters[0].item())  # the live count after the finalize (host sync)
        self.dropping = 0  # the finalize dropped the retired particles

    def retire(self, slots) -> list:
        """Hand back requests' states (each contiguous after the sort) and free their ranges."""
        return self.retire_async(slots).wait()

    def retire_async(self, slots) -> "PendingResults":
        """retire() with the bulk of the copy left running: p2c and the node charge come back on the
        arena's stream (their ranges are reused at once), pos/vel/lc on the copy stream. Those stay
        untouched until the second step from now, whose sort writes their buffer: call wait() before
        that step. All range lookups happen before any particle is marked removed, so the set is
        still sorted for them."""
        L, s = lib(), self.stream()
        torch = self.torch
        slots = list(slots)
        if not slots:
            return PendingResults([], None)
        keys = []
        for slot in slots:
            tpl, cb, nb, ob = self.res[slot]
            keys += [cb, cb + tpl.n_cells]
        k = torch.tensor(keys, dtype=torch.int32, device=self.dev)
        bounds = torch.zeros(len(keys), dtype=torch.int64, device=self.dev)
        check(L.oppx_lower_bound(self.size, self.counters.data_ptr(), self.p2c.data_ptr(), k.data_ptr(), len(keys),
                                 bounds.data_ptr(), s), "lower_bound")
        b = bounds.cpu().numpy()  # the step is complete from here on
        main = torch.cuda.current_stream(self.dev)
        self.copy_stream.wait_stream(main)
        staged = []
        pat = np.array([MAX_CELL_INDEX], np.int32)
        for j, slot in enumerate(slots):
            tpl, cb, nb, ob = self.res.pop(slot)
            lo, hi = int(b[2 * j]), int(b[2 * j + 1])
            k = hi - lo
            pin = lambda shape, dt: torch.empty(shape, dtype=dt, pin_memory=True)
            out = dict(pos=pin((3, k), torch.float64), vel=pin((3, k), torch.float64), lc=pin((4, k), torch.float64),
                       p2c=pin((k,), torch.int32), charge_f32=pin((tpl.n_nodes,), torch.float32))
            out["p2c"].copy_(self.p2c[lo:hi], non_blocking=True)
            out["charge_f32"].copy_(self.charge_f32[nb:nb + tpl.n_nodes], non_blocking=True)
            with torch.cuda.stream(self.copy_stream):
                out["pos"].copy_(self.pos[:, lo:hi], non_blocking=True)
                out["vel"].copy_(self.vel[:, lo:hi], non_blocking=True)
                out["lc"].copy_(self.lc[:, lo:hi], non_blocking=True)
            staged.append((out, k, cb))
            self.dropping += k
            if k:  # leaves the particle set at the next sort
                check(L.oppx_fill_pattern(self.p2c.data_ptr() + 4 * lo, k, 1, 4, pat.ctypes.data, 0, s),
                      "fill_pattern")
            zero = np.zeros(1, np.int32)
            check(L.oppx_fill_pattern(self.node_start.data_ptr() + 4 * nb, tpl.n_nodes, 1, 4, zero.ctypes.data, 0, s),
                  "rows")  # empty node rows
            check(L.oppx_fill_pattern(self.node_end.data_ptr() + 4 * nb, tpl.n_nodes, 1, 4, zero.ctypes.data, 0, s),
                  "rows")
            self.cells.release(cb)
            self.nodes.release(nb)
            if ob >= 0:
                self.ovl.release(ob)
            self.slots_free.append(slot)
        done = (torch.cuda.Event(), torch.cuda.Event())
        done[0].record(main)
        done[1].record(self.copy_stream)
        return PendingResults(staged, done)

    @staticmethod
    def for_budget(arena_bytes: float, device="cuda", max_insts: int = 4096) -> "FempicArena":
        """An arena of about `arena_bytes`: a quarter for mesh records (about 300 B per tet with its
        nodes, rows and overlay words), the rest for particles (about 210 B each with the sort and
        walk workspace)."""
        cells = int(0.25 * arena_bytes / 300)
        return FempicArena(max_cells=cells, max_nodes=cells // 4, max_particles=int(0.75 * arena_bytes / 210),
                           max_overlay=2 * cells, max_insts=max_insts, device=device)


class PendingResults:
    """Results of retire_async(): counts at once, the arrays after wait()."""

    def __init__(self, staged, events):
        self.staged, self.events = staged, events
        self.counts = [k for _, k, _ in staged]

    def wait(self) -> list:
        if self.events is not None:
            for e in self.events:
                e.synchronize()
            self.events = None
        return [dict(n=k, pos=o["pos"].numpy().T, vel=o["vel"].numpy().T, lc=o["lc"].numpy().T,
                     p2c=o["p2c"].numpy() - cb, charge_f32=o["charge_f32"].numpy()) for o, k, cb in self.staged]


@dataclass
class Request:
    rid: int
    arrival: float
    tpl: Template
    steps: int
    slot: int = -1
    left: int = 0
    t_admit: float = 0.0
    t_retire: float = 0.0  # ranges and budget freed
    t_done: float = 0.0    # results on the host
    result: dict = field(default_factory=dict)


def latency_split(requests) -> dict:
    """p50/p99 in ms of a request's whole latency (arrival to results on the host) and of its two
    parts: queueing (arrival to admission) and service (admission to results on the host)."""
    out = {}
    for name, f in (("latency", lambda r: r.t_done - r.arrival), ("queue", lambda r: r.t_admit - r.arrival),
                    ("service", lambda r: r.t_done - r.t_admit)):
        v = np.array([f(r) for r in requests]) * 1e3
        out[f"p50_{name}_ms"] = float(np.percentile(v, 50)) if len(v) else 0.0
        out[f"p99_{name}_ms"] = float(np.percentile(v, 99)) if len(v) else 0.0
    return out


def serve(arena: FempicArena, requests, budget_bytes: float, keep_results: bool = False, log=None):
    """Continuous batching over `requests` (sorted by arrival, seconds from start). Returns
    (requests, stats). Admission is FIFO: a request is admitted once it has arrived and fits."""
    clock = time.perf_counter
    t0 = clock()
    pending, active, done = deque(requests), {}, []
    resident = 0.0
    peak_resident, iters, part_steps = 0.0, 0, 0
    spent = dict(admit_s=0.0, step_s=0.0, retire_s=0.0)
    step_ms = []  # wall time of every batched step (all resident requests), host-synchronised
    prev = None  # (results still copying, their requests): resolved after the next step

    def finish(pr):
        pend, reqs = pr
        for r, res in zip(reqs, pend.wait()):
            r.t_done = clock() - t0
            if keep_results:
                r.result = res
            done.append(r)

    while pending or active:
        now = clock() - t0
        ta = clock()
        while pending and pending[0].arrival <= now:
            r = pending[0]
            need = r.tpl.device_bytes()
            if resident + need > budget_bytes or not arena.can_admit(r.tpl):
                break
            r.slot = arena.admit(r.tpl)
            r.left, r.t_admit = r.steps, clock() - t0
            resident += need
            peak_resident = max(peak_resident, resident)
            active[r.slot] = r
            pending.popleft()
        spent["admit_s"] += clock() - ta
        if not active:
            if prev is not None:
                finish(prev)
                prev = None
            if pending:
                time.sleep(max(0.0, min(pending[0].arrival - (clock() - t0), 0.01)))
            continue
        part_steps += arena.size - arena.dropping  # retired particles still in the set are not moved
        ts = clock()
        arena.step()  # returns once the step is complete (it reads the new live count)
        step_ms.append(1e3 * (clock() - ts))
        spent["step_s"] += step_ms[-1] * 1e-3
        iters += 1
        finished = [k for k, r in active.items() if r.left == 1]
        tr = clock()
        if prev is not None:  # their copies ran beside this step; the next step's sort reuses the buffer
            finish(prev)
            prev = None
        pend = arena.retire_async(finished)
        reqs_f = [active.pop(slot) for slot in finished]
        for r in reqs_f:
            r.t_retire = clock() - t0  # its device ranges and budget are free from here on
            resident -= r.tpl.device_bytes()
        prev = (pend, reqs_f) if reqs_f else None
        spent["retire_s"] += clock() - tr
        for r in active.values():
            r.left -= 1
    if prev is not None:
        finish(prev)
    wall = clock() - t0
    return done, dict(wall_s=wall, iterations=iters, particle_steps=part_steps, value=part_steps / wall,
                      **latency_split(done), p50_step_ms=float(np.percentile(step_ms, 50)) if step_ms else 0.0,
                      p99_step_ms=float(np.percentile(step_ms, 99)) if step_ms else 0.0,
                      peak_resident_bytes=peak_resident, requests=len(done), **spent)


def serve_plain(requests, keep_results: bool = False):
    """The plain per-request launch path: requests one at a time in arrival order, each on its own
    device state with its own launches."""
    import torch
    clock = time.perf_counter
    t0 = clock()
    part_steps = 0
    for r in requests:
        while clock() - t0 < r.arrival:
            time.sleep(max(0.0, min(r.arrival - (clock() - t0), 0.01)))
        r.t_admit = clock() - t0  # this path's admission: the request's turn on the device
        d = r.tpl.inst
        if r.tpl.proto is not None:  # a resident template: device copies of its state
            g = r.tpl.proto.clone()
        else:
            g = oppx.FempicDevice(d["mesh"], d["ef"], d["pos"], d["vel"], d["p2c"], overlay=d.get("overlay"),
                                  bmin=d.get("bmin"), bmax=d.get("bmax"), gs=d.get("gs"))
        for _ in range(r.steps):
            part_steps += g.n_max if _ == 0 else g.sync_size()
            g.step_begin()
            g.pass_fused(d["charge"], d["mass"], d["dt"], deposit=False)
            g.sort(drop_removed=True)
            g.deposit_sorted_f32()
        out = g.download()
        if keep_results:
            out["charge_f32"] = g.charge_f32.cpu().numpy()
            r.result = out
        torch.cuda.synchronize()
        r.t_done = clock() - t0
        del g
    wall = clock() - t0
    return requests, dict(wall_s=wall, particle_steps=part_steps, value=part_steps / wall, **latency_split(requests),
                          requests=len(requests))
