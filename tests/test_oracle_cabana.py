"""The oracle's CabanaPIC move with per-hop current deposit against the reference seq build's
fixtures (tests/golden/make_cabana_golden.py): positions, velocities, streaks, cells and hop counts
bit-exact, and the accumulators bit-exact too (the oracle deposits in the seq loop's particle
order). The larger case is pinned by SHA-256 digests of the particle arrays."""

import hashlib

import numpy as np
import pytest

import oracle
from conftest import golden_path


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def periodic_faces(nx, ny, nz):
    """Face neighbours {x-1, y-1, z-1, x+1, y+1, z+1} of cell x + nx*(y + ny*z), wrapped."""
    x = np.arange(nx * ny * nz)
    i, j, k = x % nx, (x // nx) % ny, x // (nx * ny)
    cid = lambda a, b, c: (a % nx) + nx * ((b % ny) + ny * (c % nz))
    return np.stack([cid(i - 1, j, k), cid(i, j - 1, k), cid(i, j, k - 1), cid(i + 1, j, k), cid(i, j + 1, k),
                     cid(i, j, k + 1)], axis=1).astype(np.int32)


def cabana_case(g):
    """Inputs of a fixture: regenerated from its seed (they are only digested in the large case)."""
    nx, ny, nz, ppc = (int(g[f"meta_{k}"]) for k in ("nx", "ny", "nz", "ppc"))
    nc = nx * ny * nz
    rng = np.random.default_rng(int(g["meta_seed"]))
    interp = rng.normal(size=(nc, 18)) * np.array([0.5, 0.2, 0.2, 0.1] * 3 + [0.5, 0.2] * 3)
    n = nc * ppc
    pos = rng.uniform(-1.0, 1.0, (n, 3))
    vel = rng.normal(0.0, float(g["meta_vel_sigma"]), (n, 3))
    weight = rng.uniform(0.5, 1.5, n)
    return dict(c2c6=periodic_faces(nx, ny, nz), interp=interp, pos=pos, vel=vel, weight=weight,
                p2c=np.repeat(np.arange(nc, dtype=np.int32), ppc), n_cells=nc)


@pytest.mark.parametrize("name", ["cabana_small.npz", "cabana_grid16.npz"])
def test_oracle_cabana_move_deposit_matches_reference(name):
    g = np.load(golden_path(name), allow_pickle=False)
    d = cabana_case(g)
    full = "in_pos" in g
    if full:
        for k in ("interp", "pos", "vel", "weight"):
            assert np.array_equal(d[k], g[f"in_{k}"]), k
    else:
        for k in ("interp", "pos", "vel", "weight"):
            assert sha(d[k]) == str(g[f"sha_in_{k}"]), k
    pos, vel, p2c = d["pos"].copy(), d["vel"].copy(), d["p2c"].copy()
    streak = np.zeros_like(pos)
    st = g["out_steps_iter_size_hops_maxhops"]
    steps = int(g["meta_steps"])
    for s in range(steps):
        acc = np.zeros((d["n_cells"], 12))
        hops = oracle.cabana_move_deposit(d["c2c6"], d["interp"], acc, pos, vel, streak, d["weight"], p2c,
                                          float(g["meta_qsp"]), float(g["meta_qdt_2mc"]), g["meta_cdt_d"])
        assert (int(hops.sum()), int(hops.max())) == tuple(st[s, 2:4]), f"step {s}"
        if s in (0, steps - 1):
            t = "s1" if s == 0 else "final"
            assert np.array_equal(acc.ravel(), g[f"out_{t}_acc"]), f"{t} acc"
            for k, a in (("pos", pos), ("vel", vel), ("streak", streak), ("p2c", p2c), ("hops", hops)):
                if full:
                    assert np.array_equal(a.ravel(), g[f"out_{t}_{k}"]), f"{t} {k}"
                else:
                    assert sha(a) == str(g[f"sha_{t}_{k}"]), f"{t} {k}"
