"""The allocation-free hole fill (oppx_holefill_noalloc, scratch in a static device array) against
oppx_holefill with a caller workspace: two identical FEM-PIC sets take the same fused move, then one
is compacted by each entry point through the C-ABI; every particle dat over the whole capacity and
the counters must be equal bit for bit. Two calls in a row, the second on a smaller set, so the
second runs over the static scratch the first one left behind."""

import ctypes
import math

import numpy as np
import pytest

import oppx

pytestmark = pytest.mark.gpu


def _pair(nx, n_part, seed):
    m = oppx.box_mesh(nx, nx, nx)
    ef = oppx.box_field_sines(m)
    sigma = 0.02
    pos, vel, p2c = oppx.box_particles(n_part, seed, sigma, m)
    bmin, bmax, gs = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1.0 / (2 * nx)
    ov = oppx.dh_overlay(m, bmin, bmax, gs)
    dt = 2.0 / nx / (sigma * math.sqrt(3.0))  # long steps: many particles leave the box
    kw = dict(overlay=ov, bmin=bmin, bmax=bmax, gs=gs)
    a = oppx.FempicDevice(m, ef, pos, vel, p2c, **kw)
    b = oppx.FempicDevice(m, ef, pos, vel, p2c, **kw)
    for g in (a, b):
        g.step_begin()
        g.pass_fused(1e-3, 1.0, dt)
    return a, b


def _state(g):
    g.torch.cuda.synchronize()
    return {k: getattr(g, k).cpu().numpy().copy() for k in ("pos", "vel", "lc", "p2c", "counters")}


@pytest.mark.parametrize("sizes", [((10, 400_000, 31), (6, 50_000, 32))], ids=["large-then-small"])
def test_holefill_noalloc_equals_holefill(sizes):
    L = oppx.lib()
    for nx, n_part, seed in sizes:
        print(f"[no-alloc] hole fill of {n_part} particles on a {nx}^3-cube box, seed {seed}", flush=True)
        a, b = _pair(nx, n_part, seed)
        a.holefill()
        rc = L.oppx_holefill_noalloc(b.n_max, b.cap, b.counters.data_ptr(), b.hole_bits.data_ptr(), b.dats, 4,
                                     b.stream())
        oppx.check(rc, "oppx_holefill_noalloc")
        sa, sb = _state(a), _state(b)
        removed = int(sa["counters"][1])
        assert 0 < removed < n_part, "the move must remove some particles for the fill to copy records"
        assert int(sa["counters"][0]) == n_part - removed
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), f"{n_part} particles: {k} differs"


def test_holefill_noalloc_rejects_oversized_set():
    L = oppx.lib()
    n = 134217728 + 1
    rc = L.oppx_holefill_noalloc(n, n, ctypes.c_void_p(8), ctypes.c_void_p(8), (oppx.DatDesc * 1)(), 1, None)
    assert rc == oppx.ERR_ARG
    assert "OPPX_HOLEFILL_NOALLOC_MAX_PARTICLES" in L.oppx_last_error().decode()
