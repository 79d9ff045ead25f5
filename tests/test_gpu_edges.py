"""Edge cases of the injection, field-side, FEM-PIC walk, CabanaPIC and NESO-grid entry points: empty
ranges are no-ops, MAX_CELL_INDEX particles are skipped, and indices out of range are counted as faults
and leave the data they would have touched unchanged (every kernel checks an index before using it)."""

import ctypes

import numpy as np
import pytest
import torch

import oppx

pytestmark = pytest.mark.gpu


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype=dtype)


def stream():
    return torch.cuda.current_stream().cuda_stream


def counters(n=0):
    c = torch.zeros(oppx.COUNTER_WORDS, dtype=torch.int32, device="cuda")
    c[0] = n
    return c


def test_inject_empty_range_and_bad_faces():
    n_faces, n_cells, cap = 4, 10, 16
    keep = [dev(np.array([1, 2, 3, 99], np.int32), torch.int32)] + [dev(np.ones((n_faces, k)), torch.float64)
                                                                  for k in (3, 3, 3, 9)]
    inlet = oppx.Inlet(n_faces, *[t.data_ptr() for t in keep])
    pos, vel = torch.zeros((3, cap), dtype=torch.float64, device="cuda"), torch.zeros((3, cap), dtype=torch.float64,
                                                                                      device="cuda")
    p2c = dev(np.array([0] * 10 + [7, 0, 3, 2, 0, 0], np.int32), torch.int32)  # 7: no such face; 3 -> cell 99
    ef = torch.zeros((n_cells, 3), dtype=torch.float64, device="cuda")
    rnd = torch.full((2, 8), 0.25, dtype=torch.float64, device="cuda")
    parts = oppx.FempicParts(cap, pos.data_ptr(), vel.data_ptr(), None, p2c.data_ptr())
    ctr = counters()
    L = oppx.lib()
    oppx.check(L.oppx_fempic_inject(10, 0, ctypes.byref(parts), ctypes.byref(inlet), n_cells, ef.data_ptr(),
                                    rnd.data_ptr(), 8, 1.0, 1.0, 1.0, 0.1, ctr.data_ptr(), stream()), "empty")
    oppx.check(L.oppx_fempic_inject(10, 4, ctypes.byref(parts), ctypes.byref(inlet), n_cells, ef.data_ptr(),
                                    rnd.data_ptr(), 8, 1.0, 1.0, 1.0, 0.1, ctr.data_ptr(), stream()), "inject")
    torch.cuda.synchronize()
    assert int(ctr[6]) == 2  # the unknown face and the face whose cell is out of range
    p = p2c.cpu().numpy()
    assert p[10] == 7 and p[12] == 3  # left as they were
    assert p[11] == 1 and p[13] == 3  # faces 0 and 2 -> cells 1 and 3
    assert torch.count_nonzero(pos[:, [10, 12]]) == 0
    # a random dat shorter than the injected range is refused before any launch
    assert L.oppx_fempic_inject(0, 9, ctypes.byref(parts), ctypes.byref(inlet), n_cells, ef.data_ptr(),
                                rnd.data_ptr(), 8, 1.0, 1.0, 1.0, 0.1, ctr.data_ptr(), stream()) != 0


def test_electric_field_bad_node_index_is_a_fault():
    c2n = dev(np.array([[0, 1, 2, 3], [0, 1, 2, 40]], np.int32), torch.int32)
    sd = torch.ones((2, 12), dtype=torch.float64, device="cuda")
    pot = torch.ones(4, dtype=torch.float64, device="cuda")
    ef = torch.zeros((2, 3), dtype=torch.float64, device="cuda")
    ctr = counters()
    oppx.check(oppx.lib().oppx_fempic_electric_field(2, ef.data_ptr(), sd.data_ptr(), c2n.data_ptr(), pot.data_ptr(),
                                                     4, ctr.data_ptr(), stream()), "efield")
    torch.cuda.synchronize()
    assert int(ctr[6]) == 1
    e = ef.cpu().numpy()
    assert np.array_equal(e[0], [-4.0, -4.0, -4.0]) and np.array_equal(e[1], [0.0, 0.0, 0.0])


def test_cabana_empty_skipped_and_bad_neighbour():
    n_cells, n = 2, 3
    c2c6 = dev(np.array([[1, 1, 1, 1, 1, 1], [0, 0, 0, 5, 0, 0]], np.int32), torch.int32)  # cell 1, x+: no cell 5
    interp = torch.zeros((n_cells, 18), dtype=torch.float64, device="cuda")
    acc = torch.zeros((n_cells, 12), dtype=torch.float64, device="cuda")
    pos = dev(np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.9, 0.0, 0.0]]).T, torch.float64)
    vel = dev(np.array([[0.1, 0.0, 0.0], [0.1, 0.0, 0.0], [5.0, 0.0, 0.0]]).T, torch.float64)
    streak = torch.zeros((3, n), dtype=torch.float64, device="cuda")
    w = torch.ones(n, dtype=torch.float64, device="cuda")
    p2c = dev(np.array([0, oppx.MAX_CELL_INDEX, 1], np.int32), torch.int32)
    hops = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    parts = oppx.CabanaParts(n, pos.data_ptr(), vel.data_ptr(), streak.data_ptr(), w.data_ptr(), p2c.data_ptr())
    mesh = oppx.CabanaMesh(n_cells, c2c6.data_ptr(), interp.data_ptr())
    cd = (ctypes.c_double * 3)(0.5, 0.5, 0.5)
    L = oppx.lib()
    ctr = counters(0)
    oppx.check(L.oppx_cabana_move_deposit(0, ctypes.byref(parts), ctypes.byref(mesh), acc.data_ptr(), -1.0, -0.01, cd,
                                          ctr.data_ptr(), hops.data_ptr(), stream()), "empty")
    ctr = counters(n)
    oppx.check(L.oppx_cabana_move_deposit(n, ctypes.byref(parts), ctypes.byref(mesh), acc.data_ptr(), -1.0, -0.01, cd,
                                          ctr.data_ptr(), hops.data_ptr(), stream()), "cabana")
    torch.cuda.synchronize()
    h, c = hops.cpu().numpy(), ctr.cpu().numpy()
    assert h[1] == 0 and p2c[1].item() == oppx.MAX_CELL_INDEX  # skipped
    assert h[0] == 1 and p2c[0].item() == 0  # moved inside its cell
    assert int(c[6]) == 1 and p2c[2].item() == 1  # crossed x+ into no cell: a fault, cell unchanged


def test_neso_grid_empty():
    m = oppx.neso_mesh(8, 8, 0.5)
    z = np.zeros((0, 2))
    g = oppx.NesoDevice(m, z, z, np.zeros(0, np.int32))
    assert g.grid
    g.step(0.1)
    assert g.read_counters()["total_hops"] == 0


@pytest.mark.parametrize("two_pass", [False, True])
def test_fempic_walk_boundary_and_corrupt_neighbour(two_pass):
    """A walk leaving through a negative c2c entry removes the particle (any negative value, as the
    reference's `opp_c2c[min_i] >= 0` test does); an entry at or past n_cells is a corrupt map,
    reported as a fault with the particle's p2c left as it was."""
    m = oppx.box_mesh(1, 1, 1)
    nc, nn = int(m["n_cells"]), int(m["n_nodes"])
    L = oppx.lib()
    outside = np.array([[1.5, 0.5, 0.5], [0.5, -0.5, 0.5], [0.5, 0.5, 2.0]])  # beyond three different faces
    inside = np.array([[0.3, 0.6, 0.45]])
    pos_h = np.vstack([outside, inside])
    n = len(pos_h)
    results = {}
    for tag, fill in (("boundary", -1), ("negative", -7), ("corrupt", nc + 3)):
        c2c = m["c2c"].copy()
        c2c[c2c < 0] = fill
        det, vol = dev(m["c_det"].reshape(-1), torch.float64), dev(m["c_vol"], torch.float64)
        c2c_d, c2n = dev(c2c.reshape(-1), torch.int32), dev(m["c2n"].reshape(-1), torch.int32)
        aux = torch.zeros(nc * ctypes.sizeof(oppx.TetAux) // 8, dtype=torch.float64, device="cuda")
        oppx.check(L.oppx_tet_aux_build(vol.data_ptr(), c2c_d.data_ptr(), nc, aux.data_ptr(), stream()), "aux")
        ef = torch.zeros(3 * nc, dtype=torch.float64, device="cuda")
        mesh = oppx.TetMesh(nc, nn, det.data_ptr(), aux.data_ptr(), c2n.data_ptr(), ef.data_ptr())
        pos = dev(pos_h.T, torch.float64)
        vel = torch.zeros((3, n), dtype=torch.float64, device="cuda")
        lc = torch.zeros((4, n), dtype=torch.float64, device="cuda")
        p2c = dev(np.zeros(n, np.int32), torch.int32)
        parts = oppx.FempicParts(n, pos.data_ptr(), vel.data_ptr(), lc.data_ptr(), p2c.data_ptr())
        holes = torch.zeros(int(L.oppx_hole_bits_words(n)), dtype=torch.int64, device="cuda")
        charge = torch.zeros(nn, dtype=torch.float64, device="cuda")
        ctr = counters(n)
        if two_pass:
            wb = int(L.oppx_fempic_work_bytes(n))
            work = torch.zeros(max(wb, 8), dtype=torch.uint8, device="cuda")
            oppx.check(L.oppx_fempic_step_begin(ctr.data_ptr(), charge.data_ptr(), nn, stream()), "begin")
            oppx.check(L.oppx_fempic_step_fused(n, ctypes.byref(parts), ctypes.byref(mesh), None, 0.0, 0.0,
                                                ctr.data_ptr(), holes.data_ptr(), work.data_ptr(), wb, None,
                                                charge.data_ptr(), stream()), "step")
        else:
            oppx.check(L.oppx_fempic_move(n, ctypes.byref(parts), ctypes.byref(mesh), None, ctr.data_ptr(),
                                          holes.data_ptr(), None, 0, None, stream()), "move")
        torch.cuda.synchronize()
        results[tag] = (p2c.cpu().numpy(), ctr.cpu().numpy(), charge.cpu().numpy())
    for tag in ("boundary", "negative"):
        p, c, q = results[tag]
        assert list(p[:3]) == [oppx.MAX_CELL_INDEX] * 3 and 0 <= p[3] < nc, tag
        assert c[1] == 3 and c[6] == 0, tag  # removed, faults
        if two_pass:
            assert abs(q.sum() - 1.0) < 1e-12, tag  # the one remaining particle deposits its weights (sum 1)
    p, c, q = results["corrupt"]
    assert c[6] == 3 and c[1] == 0
    assert p[3] == results["boundary"][0][3]
    if two_pass:
        # a walk handed to the second pass already holds the cell it stopped at
        assert ((p[:3] >= 0) & (p[:3] < nc)).all()
        assert abs(q.sum() - 1.0) < 1e-12
    else:
        assert list(p[:3]) == [0, 0, 0]
