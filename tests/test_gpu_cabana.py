"""GPU parity of the CabanaPIC move with per-hop current deposit (oppx_cabana_move_deposit) against
the reference seq build's fixtures: positions, velocities, streaks, cells and hop counts bit-exact;
the accumulators within 1e-12 relative (FP64 atomics reorder the per-cell sums)."""

import ctypes
import hashlib

import numpy as np
import pytest
import torch

import oppx
from conftest import golden_path
from test_oracle_cabana import cabana_case

pytestmark = pytest.mark.gpu
ACC_RTOL = 1e-12


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


class CabanaDevice:
    def __init__(self, d):
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype=dt)
        n = d["p2c"].size
        self.n, self.nc = n, d["n_cells"]
        self.pos, self.vel = t(d["pos"].T, torch.float64), t(d["vel"].T, torch.float64)
        self.streak = torch.zeros((3, n), dtype=torch.float64, device="cuda")
        self.weight, self.p2c = t(d["weight"], torch.float64), t(d["p2c"], torch.int32)
        self.c2c6, self.interp = t(d["c2c6"], torch.int32), t(d["interp"], torch.float64)
        self.acc = torch.zeros((self.nc, 12), dtype=torch.float64, device="cuda")
        self.hops = torch.zeros(n, dtype=torch.int32, device="cuda")
        self.ctr = torch.zeros(oppx.COUNTER_WORDS, dtype=torch.int32, device="cuda")
        self.ctr[0] = n
        self.parts = oppx.CabanaParts(n, self.pos.data_ptr(), self.vel.data_ptr(), self.streak.data_ptr(),
                                      self.weight.data_ptr(), self.p2c.data_ptr())
        self.mesh = oppx.CabanaMesh(self.nc, self.c2c6.data_ptr(), self.interp.data_ptr())

    def step(self, qsp, qdt_2mc, cdt_d):
        self.acc.zero_()
        cd = (ctypes.c_double * 3)(*[float(x) for x in cdt_d])
        oppx.check(oppx.lib().oppx_cabana_move_deposit(self.n, ctypes.byref(self.parts), ctypes.byref(self.mesh),
                                                       self.acc.data_ptr(), qsp, qdt_2mc, cd, self.ctr.data_ptr(),
                                                       self.hops.data_ptr(), torch.cuda.current_stream().cuda_stream),
                   "cabana_move_deposit")
        c = self.ctr.cpu().numpy()
        return int(c[4:6].view(np.int64)[0]), int(c[3]), int(c[6])


@pytest.mark.parametrize("name", ["cabana_small.npz", "cabana_grid16.npz"])
def test_gpu_cabana_move_deposit_matches_reference(name):
    g = np.load(golden_path(name), allow_pickle=False)
    d = cabana_case(g)
    full = "in_pos" in g
    dev = CabanaDevice(d)
    st = g["out_steps_iter_size_hops_maxhops"]
    steps = int(g["meta_steps"])
    for s in range(steps):
        tot, mx, faults = dev.step(float(g["meta_qsp"]), float(g["meta_qdt_2mc"]), g["meta_cdt_d"])
        assert faults == 0 and (tot, mx) == tuple(st[s, 2:4]), f"step {s}"
        if s in (0, steps - 1):
            t = "s1" if s == 0 else "final"
            ref = g[f"out_{t}_acc"]
            acc = dev.acc.cpu().numpy().ravel()
            assert np.abs(acc - ref).max() <= ACC_RTOL * np.abs(ref).max(), f"{t} acc"
            for k, a in (("pos", dev.pos.cpu().numpy().T), ("vel", dev.vel.cpu().numpy().T),
                         ("streak", dev.streak.cpu().numpy().T), ("p2c", dev.p2c.cpu().numpy()),
                         ("hops", dev.hops.cpu().numpy())):
                a = np.ascontiguousarray(a)
                if full:
                    assert np.array_equal(a.ravel(), g[f"out_{t}_{k}"]), f"{t} {k}"
                else:
                    assert sha(a) == str(g[f"sha_{t}_{k}"]), f"{t} {k}"
