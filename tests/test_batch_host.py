"""Host side of the batched ensemble (CPU): the concatenated mesh is the disjoint union of the
instances, so stepping the union (oracle) gives each instance's own step; index offsets and the
per-cell instance map are consistent."""

import math

import numpy as np

import oppx
import oracle


def instances(seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for i, (nx, npart) in enumerate([(2, 400), (3, 900), (4, 50), (2, 0), (3, 2000)]):
        m = oppx.box_mesh(nx, nx, nx)
        ef = oppx.box_field_sines(m)
        pos, vel, p2c = oppx.box_particles(npart, 70 + i, float(rng.uniform(0.005, 0.02)), m)
        d = dict(mesh=m, ef=ef, pos=pos, vel=vel, p2c=p2c, charge=1e-3, mass=1.0, dt=0.01)
        if i % 2 == 0:
            gs = 1.0 / (2 * nx)
            d.update(overlay=oppx.dh_overlay(m, (0, 0, 0), (1, 1, 1), gs), bmin=(0.0,) * 3, bmax=(1.0,) * 3, gs=gs)
        out.append(d)
    return out


def test_concat_offsets_and_maps():
    inst = instances()
    u = oppx.concat_instances(inst)
    assert u["mesh"]["n_cells"] == sum(d["mesh"]["n_cells"] for d in inst)
    for i, d in enumerate(inst):
        c0, n = int(u["cell_base"][i]), d["mesh"]["n_cells"]
        sl = slice(c0, c0 + n)
        assert (u["cell_inst"][sl] == i).all()
        c2c = u["mesh"]["c2c"][sl]
        assert np.array_equal(np.where(c2c >= 0, c2c - c0, c2c), d["mesh"]["c2c"])
        assert np.array_equal(u["mesh"]["c2n"][sl] - int(u["node_base"][i]), d["mesh"]["c2n"])
        assert np.array_equal(u["mesh"]["c_det"][sl], d["mesh"]["c_det"].reshape(-1, 16))
        if u["overlay_off"][i] >= 0:
            o = int(u["overlay_off"][i])
            seg = u["overlay"][o:o + d["overlay"].size]
            assert np.array_equal(np.where(seg == oppx.MAX_CELL_INDEX, seg, seg - c0), d["overlay"])
        else:
            assert "overlay" not in d


def test_union_step_is_each_instance_step():
    inst = instances(6)
    u = oppx.concat_instances(inst)
    U = oracle.FempicState(u["mesh"], u["ef"], u["pos"], u["vel"], u["p2c"])
    solo = [oracle.FempicState(d["mesh"], d["ef"], d["pos"], d["vel"], d["p2c"]) for d in inst]
    for _ in range(5):
        U.step_sorted(1e-3, 1.0, 0.01)
        for s in solo:
            s.step_sorted(1e-3, 1.0, 0.01)
    e = U.live()
    for i, s in enumerate(solo):
        c0, n = int(u["cell_base"][i]), inst[i]["mesh"]["n_cells"]
        lo, hi = np.searchsorted(e["p2c"], [c0, c0 + n])
        a = s.live()
        assert hi - lo == a["n"]
        assert np.array_equal(e["p2c"][lo:hi] - c0, a["p2c"])
        for k in ("pos", "vel", "lc"):
            assert np.array_equal(e[k][lo:hi], a[k])
        n0 = int(u["node_base"][i])
        assert np.array_equal(e["charge"][n0:n0 + inst[i]["mesh"]["n_nodes"]], a["charge"])


def test_concat_rejects_bad_input():
    inst = instances()
    bad = dict(inst[0])
    bad["p2c"] = bad["p2c"].copy()
    if len(bad["p2c"]):
        bad["p2c"][0] = bad["mesh"]["n_cells"]
        try:
            oppx.concat_instances([bad])
            raise AssertionError("accepted p2c out of range")
        except ValueError:
            pass
