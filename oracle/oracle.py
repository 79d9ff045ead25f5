"""ctypes binding of the oracle (liboracle.so): the CPU checker of the GPU hot path.

Test infrastructure only: imported by tests/, the smoke check and the CPU-baseline leg of
bench.py, never by the product.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
MAX_CELL_INDEX = 2147483647

_SIGS = {
    "orc_fempic_push": (None, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_double, c_int]),
    "orc_fempic_move": (c_int64, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "orc_fempic_deposit": (None, [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "orc_holefill": (c_int64, [c_int64, c_int64, c_void_p, POINTER(c_void_p), c_void_p, c_void_p, c_int]),
    "orc_sort_by_cell": (None, [c_int64, c_void_p, POINTER(c_void_p), c_void_p, c_void_p, c_int, c_void_p]),
    "orc_fempic_step": (c_int64, [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_double, c_double, c_double, c_int, c_void_p]),
    "orc_dh_overlay_bruteforce": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                                          c_void_p, c_int]),
    "orc_neso_step": (c_int64, [c_int64, c_int, c_int, c_double, c_double, c_double, c_double, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "orc_inc_part_distribution": (None, [c_int64, c_int64, c_void_p, c_int, c_void_p]),
    "orc_fempic_inject": (None, [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_double, c_double]),
    "orc_node_charge_density": (None, [c_int64, c_void_p, c_void_p, c_double]),
    "orc_electric_field": (None, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "orc_boundary_potential": (None, [c_int64, c_void_p, c_double, c_void_p]),
    "orc_max_fold": (c_double, [c_int64, c_void_p, c_int, c_double]),
    "orc_cabana_move_deposit": (c_int64, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_double, c_double, c_void_p, c_void_p]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build with `make -C oracle oracle`")
        L = ctypes.CDLL(LIB_PATH)
        for n, (r, a) in _SIGS.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


class FempicState:
    """Host (AoS) state of one FEM-PIC instance, stepped by the oracle."""

    def __init__(self, mesh, c_ef, pos, vel, p2c, overlay=None, bmin=None, gs=None, bmax=None):
        self.c2n = np.ascontiguousarray(mesh["c2n"], np.int32)
        self.c2c = np.ascontiguousarray(mesh["c2c"], np.int32)
        self.c_det = np.ascontiguousarray(mesh["c_det"], np.float64)
        self.c_vol = np.ascontiguousarray(mesh["c_vol"], np.float64)
        self.c_ef = np.ascontiguousarray(c_ef, np.float64)
        self.nn = int(mesh["n_nodes"])
        self.pos = np.array(pos, np.float64, copy=True)
        self.vel = np.array(vel, np.float64, copy=True)
        self.lc = np.zeros((pos.shape[0], 4))
        self.p2c = np.array(p2c, np.int32, copy=True)
        self.n = int(pos.shape[0])
        self.charge = np.zeros(self.nn)
        self.overlay = None if overlay is None else np.ascontiguousarray(overlay, np.int32)
        if overlay is not None:
            d = []
            for k in range(3):
                a, x = 0, float(bmin[k])
                while x < bmax[k]:
                    a += 1
                    x += gs
                d.append(a + 1)
            self.dims = np.array(d, np.int64)
            self.bmin = np.ascontiguousarray(bmin, np.float64)
            self.inv_gs = 1.0 / gs
        else:
            self.dims = np.zeros(3, np.int64)
            self.bmin = np.zeros(3)
            self.inv_gs = 0.0

    def step(self, charge, mass, dt, threads=1):
        stats = np.zeros(4, np.int64)
        self.n = int(lib().orc_fempic_step(self.n, self.nn, _p(self.c2n), _p(self.c2c), _p(self.c_det),
                                           _p(self.c_vol), _p(self.c_ef), _p(self.overlay), _p(self.dims),
                                           _p(self.bmin), self.inv_gs, _p(self.pos), _p(self.vel), _p(self.lc),
                                           _p(self.p2c), _p(self.charge), charge, mass, dt, threads, _p(stats)))
        return dict(removed=int(stats[0]), total_hops=int(stats[1]), max_hops=int(stats[2]), size=self.n)

    def push(self, charge, mass, dt, threads=1):
        lib().orc_fempic_push(self.n, _p(self.c_ef), _p(self.p2c), _p(self.pos), _p(self.vel), charge, mass, dt,
                              threads)

    def move(self, threads=1):
        hops = np.zeros(self.n, np.int32)
        removed = lib().orc_fempic_move(self.n, _p(self.c2c), _p(self.c_det), _p(self.c_vol), _p(self.overlay),
                                        _p(self.dims), _p(self.bmin), self.inv_gs, _p(self.pos), _p(self.lc),
                                        _p(self.p2c), _p(hops), threads)
        return int(removed), hops

    def holefill(self, removed):
        marks = self.p2c[: self.n].copy()
        arrs = [self.pos, self.vel, self.lc, self.p2c]
        ptrs = (c_void_p * 4)(*[a.ctypes.data for a in arrs])
        dims = np.array([3, 3, 4, 1], np.int32)
        eb = np.array([8, 8, 8, 4], np.int32)
        self.n = int(lib().orc_holefill(self.n, removed, _p(marks), ptrs, _p(dims), _p(eb), 4))

    def deposit(self, threads=1):
        lib().orc_fempic_deposit(self.n, self.nn, _p(self.c2n), _p(self.p2c), _p(self.lc), _p(self.charge), threads)

    def sort(self, drop_removed=True):
        """Stable sort by (p2c, index) with removed last; drop_removed shrinks n by the removed count."""
        n = self.n
        removed = int((self.p2c[:n] == MAX_CELL_INDEX).sum())
        arrs = [self.pos[:n].copy(), self.vel[:n].copy(), self.lc[:n].copy(), self.p2c[:n].copy()]
        perm = sort_by_cell(n, arrs, self.p2c)
        self.pos[:n], self.vel[:n], self.lc[:n], self.p2c[:n] = arrs
        if drop_removed:
            self.n = n - removed
        return perm

    def step_sorted(self, charge, mass, dt, threads=1):
        """The step with the Sort_All finalize: push, move, sort (removed last, dropped), deposit."""
        self.charge[:] = 0.0
        self.push(charge, mass, dt, threads)
        removed, hops = self.move(threads)
        self.sort(drop_removed=True)
        self.deposit(threads)
        return dict(removed=removed, total_hops=int(hops.sum()), max_hops=int(hops.max()) if hops.size else 0,
                    size=self.n)

    def live(self):
        n = self.n
        return dict(n=n, pos=self.pos[:n].copy(), vel=self.vel[:n].copy(), lc=self.lc[:n].copy(),
                    p2c=self.p2c[:n].copy(), charge=self.charge.copy())


class FempicLoop(FempicState):
    """The FEM-PIC main loop minus the field solve (apps/app_fempic/fempic.cpp:138-231): injection
    through an inlet, push, move, hole fill, deposit, node charge density, electric field from a
    given node potential, and the loop's OPP_MAX globals. Starts with no particles."""

    def __init__(self, mesh, c_ef, sd, n_vol, n_type, n_pot, inlet, rand, capacity, overlay=None, bmin=None,
                 gs=None, bmax=None):
        z = np.zeros((capacity, 3))
        super().__init__(mesh, c_ef, z, z, np.zeros(capacity, np.int32), overlay=overlay, bmin=bmin, gs=gs,
                         bmax=bmax)
        self.n, self.diff = 0, 0
        self.nc = int(mesh["n_cells"])
        self.sd = np.ascontiguousarray(sd, np.float64)
        self.n_vol = np.ascontiguousarray(n_vol, np.float64)
        self.n_type = np.ascontiguousarray(n_type, np.int32)
        self.n_pot = np.ascontiguousarray(n_pot, np.float64)
        self.inlet = {k: np.ascontiguousarray(inlet[k]) for k in ("if2c", "if_u", "if_v", "if_norm", "if_npos",
                                                                  "if_dist")}
        self.rand = np.ascontiguousarray(rand, np.float64)

    def inc_distribution(self, count):
        """opp_inc_part_count_with_distribution: count new particles holding their inlet face."""
        if self.n + count > self.p2c.shape[0]:
            raise ValueError("capacity exceeded")
        d = self.inlet["if_dist"]
        lib().orc_inc_part_distribution(self.n, count, _p(d), d.shape[0], _p(self.p2c))
        self.n += count
        self.diff = count

    def inject(self, ion_velocity, charge, mass, dt):
        i = self.inlet
        lib().orc_fempic_inject(self.n - self.diff, self.diff, _p(i["if2c"]), _p(self.c_ef), _p(i["if_u"]),
                                _p(i["if_v"]), _p(i["if_norm"]), _p(i["if_npos"]), _p(self.rand), _p(self.pos),
                                _p(self.vel), _p(self.p2c), ion_velocity, charge, mass, dt)

    def boundary_potential(self, wall):
        out = np.zeros(self.nn)
        lib().orc_boundary_potential(self.nn, _p(self.n_type), wall, _p(out))
        return out

    def field_side(self, spwt):
        """compute_node_charge_density, then c_ef reset and compute_electric_field; returns the
        loop's maxima (max c_ef component, max |charge density|, max |potential|)."""
        lib().orc_node_charge_density(self.nn, _p(self.charge), _p(self.n_vol), spwt)
        self.c_ef[:] = 0.0
        lib().orc_electric_field(self.nc, _p(self.c_ef), _p(self.sd), _p(self.c2n), _p(self.n_pot))
        L = lib()
        return (L.orc_max_fold(3 * self.nc, _p(self.c_ef), 0, 0.0), L.orc_max_fold(self.nn, _p(self.charge), 1, 0.0),
                L.orc_max_fold(self.nn, _p(self.n_pot), 1, 0.0))

    def loop_step(self, inject_count, first, ion_velocity, charge, mass, dt, spwt):
        """One main-loop iteration; returns (size before the move, size after, total hops, max hops,
        maxima)."""
        if not first:
            self.inc_distribution(inject_count)
        before = self.n
        self.inject(ion_velocity, charge, mass, dt)
        self.charge[:] = 0.0
        self.push(charge, mass, dt)
        removed, hops = self.move()
        self.holefill(removed)
        self.deposit()
        mx = self.field_side(spwt)
        return before, self.n, int(hops.sum()), int(hops.max()) if hops.size else 0, mx


def cabana_move_deposit(c2c6, interp, acc, pos, vel, streak, weight, p2c, qsp, qdt_2mc, cdt_d):
    """CabanaPIC move_deposit over every particle (AoS arrays, updated in place); returns the
    per-particle evaluation counts."""
    n = p2c.shape[0]
    hops = np.zeros(n, np.int32)
    cd = np.ascontiguousarray(cdt_d, np.float64)
    lib().orc_cabana_move_deposit(n, _p(c2c6), _p(interp), _p(acc), _p(pos), _p(vel), _p(streak), _p(weight),
                                  _p(p2c), qsp, qdt_2mc, _p(cd), _p(hops))
    return hops


def dh_overlay_bruteforce(mesh, bmin, bmax, gs, threads=0):
    bmin = np.ascontiguousarray(bmin, np.float64)
    bmax = np.ascontiguousarray(bmax, np.float64)
    d = []
    for k in range(3):
        a, x = 0, float(bmin[k])
        while x < bmax[k]:
            a += 1
            x += gs
        d.append(a + 1)
    ov = np.zeros(int(np.prod(d)), np.int32)
    lib().orc_dh_overlay_bruteforce(mesh["n_cells"], _p(np.ascontiguousarray(mesh["c2c"], np.int32)),
                                    _p(np.ascontiguousarray(mesh["c_det"])), _p(np.ascontiguousarray(mesh["c_vol"])),
                                    _p(bmin), _p(bmax), gs, _p(ov), threads)
    return ov


def sort_by_cell(n, arrays, p2c):
    """Stable by (p2c, index); arrays = list of AoS numpy arrays (p2c included), permuted in place."""
    ptrs = (c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
    dims = np.array([1 if a.ndim == 1 else a.shape[1] for a in arrays], np.int32)
    eb = np.array([a.itemsize for a in arrays], np.int32)
    perm = np.zeros(n, np.int64)
    keys = np.ascontiguousarray(p2c[:n], np.int32).copy()
    lib().orc_sort_by_cell(n, _p(keys), ptrs, _p(dims), _p(eb), len(arrays), _p(perm))
    return perm


def neso_step(n, nx, ny, cell_width, dt, c2c, c_pos_ll, pos, vel, mdir, p2c, threads=1):
    hops = np.zeros(n, np.int32)
    tot = lib().orc_neso_step(n, nx, ny, cell_width, dt, nx * cell_width, ny * cell_width, _p(c2c), _p(c_pos_ll),
                              _p(pos), _p(vel), _p(mdir), _p(p2c), _p(hops), threads)
    return int(tot), hops
