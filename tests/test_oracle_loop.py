"""The oracle's FEM-PIC main loop minus the field solve (injection, push, move, hole fill, deposit,
node charge density, electric field, OPP_MAX globals) against the reference seq build's fixtures
(tests/golden/make_fempic_loop_golden.py): every array bit-exact, the maxima exact. This pins the
oracle's restatement of inject_ions, the distribution fill, compute_node_charge_density,
compute_electric_field, init_boundary_pot and the max folds, including the FMA contractions of
the reference's GCC build. The third case runs on a 12-tet box read through the mesh-file loader, with
the loader's node volumes, node types and inlet faces (the reference app's own mesh input path)."""

import numpy as np
import pytest

import oppx
import oracle
from conftest import golden_path


def loop_mesh(g):
    """The fixture's mesh: the Kuhn box, or the 12-tet box read through the mesh-file loader."""
    if "meta_mesh12" in g:
        nx, ny, nz = (int(v) for v in g["meta_mesh12"])
        m = oppx.box12_mesh(nx, ny, nz, 1.0, ny / nx, nz / nx)
        bmax = tuple(float(v) for v in g["meta_bmax"])
    else:
        nx = int(g["meta_nx"])
        m = oppx.box_mesh(nx, nx, nx)
        bmax = (1.0, 1.0, 1.0)
    return m, oppx.box_field_sines(m), bmax


def loop_case(name):
    g = np.load(golden_path(name), allow_pickle=False)
    m, ef, bmax = loop_mesh(g)
    inlet = {k: g[f"in_{k}"] for k in ("if2c", "if_u", "if_v", "if_norm", "if_npos", "if_dist")}
    inject, steps = int(g["meta_inject"]), int(g["meta_steps"])
    kw = {}
    if int(g["meta_dh"]):
        gs = float(g["meta_gs"])
        kw = dict(overlay=oppx.dh_overlay(m, (0, 0, 0), bmax, gs), bmin=(0, 0, 0), bmax=bmax, gs=gs)
    o = oracle.FempicLoop(m, ef, g["in_sd"], g["in_n_vol"], g["in_n_type"], g["in_n_pot"], inlet, g["in_rand"],
                          capacity=inject * (steps + 1), **kw)
    return g, o, inject, steps


@pytest.mark.parametrize("name", ["fempic_loop_small.npz", "fempic_loop_dh.npz", "fempic_loop_mesh12.npz"])
def test_oracle_main_loop_matches_reference(name):
    g, o, inject, steps = loop_case(name)
    q, ms, dt = float(g["meta_charge"]), float(g["meta_mass"]), float(g["meta_dt"])
    iv, spwt = float(g["meta_ion_velocity"]), float(g["meta_spwt"])
    assert np.array_equal(o.boundary_potential(float(g["meta_wall_potential"])), g["out_bnd_pot"])
    o.inc_distribution(inject)
    assert np.array_equal(o.p2c[: o.n], g["out_dist1_p2c"])
    st, mx = g["out_steps_iter_before_after_hops_maxhops"], g["out_steps_max_cef_chg_pot"]
    for s in range(steps):
        if s == 0:
            o.inject(iv, q, ms, dt)
            n = o.n
            assert np.array_equal(o.pos[:n].ravel(), g["out_inj1_pos"])
            assert np.array_equal(o.vel[:n].ravel(), g["out_inj1_vel"])
            assert np.array_equal(o.p2c[:n], g["out_inj1_p2c"])
            o.diff = 0  # injected already: the step below must not inject them again
        before, after, tot, mh, maxima = o.loop_step(inject, s == 0, iv, q, ms, dt, spwt)
        if s == 0:
            before = n
        assert (before, after, tot, mh) == tuple(st[s, 1:5]), f"step {s}"
        assert np.array_equal(np.array(maxima), mx[s]), f"step {s} maxima"
        if s in (0, steps - 1):
            t = "s1" if s == 0 else "final"
            n = o.n
            for k, a in (("pos", o.pos[:n]), ("vel", o.vel[:n]), ("lc", o.lc[:n]), ("p2c", o.p2c[:n]),
                         ("charge", o.charge), ("ef", o.c_ef)):
                assert np.array_equal(a.ravel(), g[f"out_{t}_{k}"]), f"{t} {k}"


def test_max_fold_keeps_the_first_maximum():
    # the MAX macro keeps m unless v is strictly greater: a NaN never enters, and -0.0 after +0.0
    # leaves +0.0 (closed forms of the fold)
    x = np.array([-0.0, np.nan, -3.0, 0.0, -0.0])
    assert np.signbit(oracle.lib().orc_max_fold(5, oracle._p(x), 0, -0.0))
    y = np.array([np.nan, -1.0, 2.0, 2.0, np.nan])
    assert oracle.lib().orc_max_fold(5, oracle._p(y), 0, 0.0) == 2.0
    assert oracle.lib().orc_max_fold(5, oracle._p(-y), 1, 0.0) == 2.0
