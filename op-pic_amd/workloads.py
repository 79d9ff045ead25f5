"""Seeded synthetic FEM-PIC instances: the box mesh (n^3 cubes x 6 tets), the sin-product field,
uniform particles with Maxwellian velocities and a direct-hop overlay, as one dict per instance
(the input format of FempicDevice / FempicBatch / the serving loop)."""

from __future__ import annotations

import math

import numpy as np

import oppx


def box_instance(nx: int, n_part: int, seed: int, sigma: float, overlay_ratio: float = 2.0,
                 charge: float = 1e-3, mass: float = 1.0) -> dict:
    """One instance on [0,1]^3: dt = 0.5 h / (sigma sqrt 3) (h = 1/nx), overlay spacing
    1 / round(overlay_ratio * nx)."""
    m = oppx.box_mesh(nx, nx, nx)
    ef = oppx.box_field_sines(m)
    pos, vel, p2c = oppx.box_particles(n_part, seed, sigma, m)
    pos, vel, p2c = oppx.sort_particles_by_cell(pos, vel, p2c)
    gs = 1.0 / max(1, int(round(overlay_ratio * nx)))
    bmin, bmax = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    return dict(mesh=m, ef=ef, pos=pos, vel=vel, p2c=p2c, charge=charge, mass=mass,
                dt=0.5 * (1.0 / nx) / (sigma * math.sqrt(3.0)), overlay=oppx.dh_overlay(m, bmin, bmax, gs),
                bmin=bmin, bmax=bmax, gs=gs, nx=nx, sigma=sigma, seed=seed)


def loguniform(rng, lo, hi):
    return float(math.exp(rng.uniform(math.log(lo), math.log(hi))))


def ensemble_specs(n_inst: int = 256, seed: int = 4, tets=(2_000, 60_000), parts=(10_000, 2_000_000),
                   sigma=(0.005, 0.02), particle_scale: float = 1.0):
    """(nx, n_part, seed, sigma) of each instance: tets and particles log-uniform, sigma uniform.
    A box of nx^3 cubes has 6 nx^3 tets, so nx = round((tets / 6)^(1/3))."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_inst):
        nx = max(2, int(round((loguniform(rng, *tets) / 6.0) ** (1.0 / 3.0))))
        n_part = max(1, int(round(loguniform(rng, *parts) * particle_scale)))
        out.append((nx, n_part, seed * 100_003 + i, float(rng.uniform(*sigma))))
    return out


def ensemble(n_inst: int = 256, seed: int = 4, particle_scale: float = 1.0, **kw):
    return [box_instance(nx, n, s, sg) for (nx, n, s, sg) in ensemble_specs(n_inst, seed,
                                                                              particle_scale=particle_scale, **kw)]
