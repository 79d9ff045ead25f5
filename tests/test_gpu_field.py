"""GPU kernels of the injection and field-side loops through the C-ABI, against the oracle on seeded
inputs larger than the reference fixtures cover: inject_ions, compute_node_charge_density,
compute_electric_field and init_boundary_pot bit-exact; the OPP_MAX fold exact, including which of
+0.0 / -0.0 it keeps and that NaNs never enter, over many blocks."""

import ctypes

import numpy as np
import pytest
import torch

import oppx
import oracle

pytestmark = pytest.mark.gpu


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype=dtype)


def stream():
    return torch.cuda.current_stream().cuda_stream


def test_gpu_inject_matches_oracle():
    rng = np.random.default_rng(41)
    n_faces, n_cells, cap, first, count = 300, 5000, 70000, 1234, 60000
    if2c = rng.integers(0, n_cells, n_faces).astype(np.int32)
    u, v, nrm = rng.normal(size=(n_faces, 3)), rng.normal(size=(n_faces, 3)), rng.normal(size=(n_faces, 3))
    npos = rng.normal(size=(n_faces, 9))
    ef = rng.normal(size=(n_cells, 3))
    rnd = rng.random((count + 100, 2))
    p2c = np.full(cap, 7, np.int32)
    p2c[first:first + count] = rng.integers(0, n_faces, count)
    pos, vel = rng.normal(size=(cap, 3)), rng.normal(size=(cap, 3))
    iv, q, m, dt = 0.37, 1.6e-3, 2.5, 0.011
    # oracle (AoS)
    o_pos, o_vel, o_p2c = pos.copy(), vel.copy(), p2c.copy()
    oracle.lib().orc_fempic_inject(first, count, oracle._p(if2c), oracle._p(ef), oracle._p(u), oracle._p(v),
                                   oracle._p(nrm), oracle._p(npos), oracle._p(rnd), oracle._p(o_pos),
                                   oracle._p(o_vel), oracle._p(o_p2c), iv, q, m, dt)
    # device (SoA particles, SoA random pairs)
    d_pos, d_vel, d_p2c = dev(pos.T, torch.float64), dev(vel.T, torch.float64), dev(p2c, torch.int32)
    d_rnd = dev(rnd.T, torch.float64)
    keep = [dev(if2c, torch.int32), dev(u, torch.float64), dev(v, torch.float64), dev(nrm, torch.float64),
            dev(npos, torch.float64)]
    inlet = oppx.Inlet(n_faces, *[t.data_ptr() for t in keep])
    parts = oppx.FempicParts(cap, d_pos.data_ptr(), d_vel.data_ptr(), None, d_p2c.data_ptr())
    d_ef = dev(ef, torch.float64)
    ctr = torch.zeros(oppx.COUNTER_WORDS, dtype=torch.int32, device="cuda")
    oppx.check(oppx.lib().oppx_fempic_inject(first, count, ctypes.byref(parts), ctypes.byref(inlet), n_cells,
                                             d_ef.data_ptr(), d_rnd.data_ptr(), rnd.shape[0], iv, q, m, dt,
                                             ctr.data_ptr(), stream()), "inject")
    torch.cuda.synchronize()
    assert int(ctr[6]) == 0
    assert np.array_equal(d_pos.cpu().numpy().T, o_pos)
    assert np.array_equal(d_vel.cpu().numpy().T, o_vel)
    assert np.array_equal(d_p2c.cpu().numpy(), o_p2c)


def test_gpu_field_side_loops_match_oracle():
    rng = np.random.default_rng(42)
    n_nodes, n_cells = 20000, 90000
    c2n = rng.integers(0, n_nodes, (n_cells, 4)).astype(np.int32)
    sd, pot = rng.normal(size=(n_cells, 12)), rng.normal(size=n_nodes)
    ef0 = rng.normal(size=(n_cells, 3))
    q, vol = rng.normal(size=n_nodes), rng.random(n_nodes) + 0.1
    ntype = rng.integers(0, 4, n_nodes).astype(np.int32)
    o_ef, o_q, o_b = ef0.copy(), q.copy(), np.zeros(n_nodes)
    L = oracle.lib()
    L.orc_electric_field(n_cells, oracle._p(o_ef), oracle._p(sd), oracle._p(c2n), oracle._p(pot))
    L.orc_node_charge_density(n_nodes, oracle._p(o_q), oracle._p(vol), 2.5)
    L.orc_boundary_potential(n_nodes, oracle._p(ntype), -0.75, oracle._p(o_b))
    d_ef, d_q, d_b = dev(ef0, torch.float64), dev(q, torch.float64), torch.zeros(n_nodes, dtype=torch.float64,
                                                                                 device="cuda")
    keep = [dev(sd, torch.float64), dev(c2n, torch.int32), dev(pot, torch.float64), dev(vol, torch.float64),
            dev(ntype, torch.int32)]
    X = oppx.lib()
    oppx.check(X.oppx_fempic_electric_field(n_cells, d_ef.data_ptr(), keep[0].data_ptr(), keep[1].data_ptr(),
                                            keep[2].data_ptr(), n_nodes, None, stream()), "efield")
    oppx.check(X.oppx_fempic_node_charge_density(n_nodes, d_q.data_ptr(), keep[3].data_ptr(), 2.5, stream()),
               "density")
    oppx.check(X.oppx_fempic_boundary_potential(n_nodes, keep[4].data_ptr(), -0.75, d_b.data_ptr(), stream()),
               "boundary")
    torch.cuda.synchronize()
    assert np.array_equal(d_ef.cpu().numpy(), o_ef)
    assert np.array_equal(d_q.cpu().numpy(), o_q)
    b = d_b.cpu().numpy()
    assert np.array_equal(b, o_b) and np.array_equal(np.signbit(b), np.signbit(o_b))


@pytest.mark.parametrize("n,case", [(1, "zeros"), (5000, "zeros"), (3_000_000, "zeros"), (3_000_000, "random"),
                                    (2049, "nan"), (0, "random")])
def test_gpu_max_reduce_is_the_ordered_fold(n, case):
    rng = np.random.default_rng(n)
    if case == "zeros":  # all maxima tie at zero: which sign is kept depends on the order alone
        x = np.where(rng.random(n) < 0.5, -0.0, 0.0) - (rng.random(n) < 0.3)
    elif case == "nan":
        x = rng.normal(size=n)
        x[rng.random(n) < 0.2] = np.nan
    else:
        x = rng.normal(size=n)
    d = dev(x, torch.float64)
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    nb = int(oppx.lib().oppx_max_scratch_bytes(n))
    scratch = torch.zeros(max(nb, 16), dtype=torch.uint8, device="cuda")
    for absv in (0, 1):
        for init in (0.0, -0.0, -np.inf):
            oppx.check(oppx.lib().oppx_max_reduce(n, d.data_ptr(), absv, init, out.data_ptr(), scratch.data_ptr(),
                                                  scratch.numel(), stream()), "max_reduce")
            got = float(out.item())
            exp = oracle.lib().orc_max_fold(n, oracle._p(x), absv, init)
            assert got == exp and np.signbit(got) == np.signbit(exp), (absv, init, got, exp)
