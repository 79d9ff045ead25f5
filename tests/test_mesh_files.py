"""FEM-PIC text mesh files (mesh.dat / inlet.dat / wall.dat): the 12-tet box writer and the loader.

The loader is pinned against the reference app's own loaders (LoadVolumeMesh / LoadSurfaceMesh of
apps/app_fempic/minifempic_funcs.h, built from the reference sources; tests/golden/mesh12_loader.npz):
bit for bit on a mesh with jittered node coordinates (so that every rounding of the geometry matters),
and by digests on the 20 x 20 x 10-cube, 48,000-tet box. The writer is pinned by what the 12-tet split
fixes: counts, positive volumes summing to the box, symmetric face neighbours, boundary faces, inlet
faces, node volumes and types. The overlay builder is checked against the reference's overlay on a
12-tet box. CPU only."""

import os
import tempfile

import numpy as np
import pytest

import oppx
from conftest import golden_path

KEYS = ["node_pos", "node_vol", "node_type", "c2n", "c2c", "c_det", "c_vol", "if2c", "if2n", "if_u", "if_v",
        "if_normal", "if_area"]


def test_loader_matches_reference_loader_on_jittered_mesh():
    g = np.load(golden_path("mesh12_loader.npz"), allow_pickle=False)
    d = golden_path("mesh12_jitter")
    m = oppx.mesh_from_files(os.path.join(d, "mesh.dat"), os.path.join(d, "inlet.dat"), os.path.join(d, "wall.dat"))
    for k in KEYS:
        assert np.array_equal(m[k].ravel(), g["jitter_" + k]), k


@pytest.mark.slow
def test_loader_matches_reference_loader_on_box48k():
    import hashlib

    g = np.load(golden_path("mesh12_loader.npz"), allow_pickle=False)
    with tempfile.TemporaryDirectory() as td:
        oppx.box12_write(20, 20, 10, 50.0, td)
        m = oppx.mesh_from_files(f"{td}/mesh.dat", f"{td}/inlet.dat", f"{td}/wall.dat")
    assert [m["n_nodes"], m["n_cells"], m["n_ifaces"]] == g["box48k_sizes"].tolist()
    for k in KEYS:
        assert hashlib.sha256(np.ascontiguousarray(m[k]).tobytes()).hexdigest() == str(g["box48k_sha_" + k]), k


@pytest.mark.parametrize("nx,ny,nz", [(1, 1, 1), (3, 2, 4), (5, 5, 2)])
def test_box12_writer_closed_forms(nx, ny, nz):
    L = 1.0 / nx
    m = oppx.box12_mesh(nx, ny, nz, 1.0, ny * L, nz * L)
    ncube = nx * ny * nz
    assert m["n_nodes"] == (nx + 1) * (ny + 1) * (nz + 1) + ncube and m["n_cells"] == 12 * ncube
    assert m["n_ifaces"] == 2 * nx * ny
    vol = 1.0 * (ny * L) * (nz * L)
    assert (m["c_vol"] > 0).all()
    assert abs(m["c_vol"].sum() - vol) <= 1e-12 * vol
    # every tet of cube q holds q's centre node, the cube's 12 tets fill it
    centre = (nx + 1) * (ny + 1) * (nz + 1) + np.repeat(np.arange(ncube), 12)
    assert (m["c2n"] == centre[:, None]).any(axis=1).all()
    assert np.allclose(m["c_vol"].reshape(ncube, 12).sum(axis=1), L ** 3, rtol=1e-12)
    # face neighbours are symmetric, boundary faces are the box surface split in two per cube face
    c2c = m["c2c"]
    for c in range(m["n_cells"]):
        for v in range(4):
            d = c2c[c, v]
            if d >= 0:
                assert c in c2c[d]
    assert (c2c < 0).sum() == 2 * 2 * (nx * ny + ny * nz + nx * nz)
    # node volumes partition the box; types: inlet z = 0, fixed side walls off the end planes, open top
    assert abs(m["node_vol"].sum() - vol) <= 1e-12 * vol
    z = m["node_pos"][:, 2]
    assert (m["node_type"][np.isclose(z, 0.0)] == 2).all()
    side = (~np.isclose(z, 0.0)) & (~np.isclose(z, nz * L)) & (
        np.isclose(m["node_pos"][:, 0], 0) | np.isclose(m["node_pos"][:, 0], 1.0) |
        np.isclose(m["node_pos"][:, 1], 0) | np.isclose(m["node_pos"][:, 1], ny * L))
    assert (m["node_type"][side] == 3).all()
    # inlet faces: in the z = 0 plane, owned by a tet holding all three nodes, normal along z, area of half a face
    assert np.allclose(m["node_pos"][m["if2n"].ravel(), 2], 0.0)
    for f in range(m["n_ifaces"]):
        assert set(m["if2n"][f]) <= set(m["c2n"][m["if2c"][f]])
    assert np.allclose(np.abs(m["if_normal"][:, 2]), 1.0)
    assert np.allclose(m["if_area"], 0.5 * L * L, rtol=1e-12)


def test_box12_particles_in_their_tets():
    m = oppx.box12_mesh(3, 3, 2, 1.0, 1.0, 2.0 / 3.0)
    pos, vel, p2c = oppx.box12_particles(5000, 9, 0.01, m)
    # barycentric weights of every particle in its tet lie in [-1e-10, 1 + 1e-10]
    d = m["c_det"][p2c].reshape(-1, 4, 4)
    lc = (d[:, :, 0] - pos[:, None, 0] * d[:, :, 1] + pos[:, None, 1] * d[:, :, 2] - pos[:, None, 2] * d[:, :, 3])
    lc = lc / (6.0 * m["c_vol"][p2c])[:, None]
    assert (lc > -1e-9).all() and (lc < 1 + 1e-9).all()
    assert np.allclose(lc.sum(axis=1), 1.0)


def test_overlay_builder_matches_reference_overlay_on_12_tet_box():
    g = np.load(golden_path("fempic_mesh12_dh.npz"), allow_pickle=False)
    nx, nz = int(g["meta_nx"]), int(g["meta_nz"])
    m = oppx.box12_mesh(nx, int(g["meta_ny"]), nz, 1.0, 1.0, nz / nx)
    ov = oppx.dh_overlay(m, tuple(g["bmin"]), tuple(g["bmax"]), float(g["meta_gs"]))
    assert np.array_equal(ov, g["out_overlay"])
