"""OP-PIC HDF5 mesh files: the library's HDF5 reader (oppx_h5_dataset_info / oppx_h5_read, behind the
drop-in's opp_decl_*_hdf5) and the reference's own shipped meshes (scripts/batch/fempic/mesh_files_tar:
box_24000, box_48000, box_96000; digests in tests/golden/fempic_h5_meshes.json, made by
tests/golden/make_h5_mesh_digests.py with an independent pure-Python reader).

Pinned: every dataset the C++ reader returns from the reference's files has the digest the pure-Python
reader gives; the 12-tet box writer reproduces the three meshes' nodes, connectivity, face neighbours,
node types and inlet faces bit for bit; their stored geometry (c_det, c_volume, n_vol) is bit for bit
what the reference's own loader (apps/app_fempic/minifempic_funcs.h) computes from the writer's text
files when compiled without FMA contraction, and within 1e-12 relative of the library's loader (which
follows the reference's contracting -O3 -march build). Error paths and type conversions on files
written by tests/h5write.py. CPU only; the reference-file tests are skipped where the reference tree
or its reference build is absent."""

import hashlib
import json
import os
import subprocess
import tarfile
import tempfile

import numpy as np
import pytest

import h5write
import oppx
from conftest import golden_path

REF = "/root/reference"
TARS = os.path.join(REF, "scripts/batch/fempic/mesh_files_tar")
MESH_REF_NOFMA = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "mesh_ref_nofma")
MESHES = json.load(open(golden_path("fempic_h5_meshes.json")))
NAMES = sorted(MESHES)


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def box_files(name, td):
    M = MESHES[name]
    nx, ny, nz = M["dims"]
    oppx.box12_write(nx, ny, nz, M["cube_mm"], td)
    return [os.path.join(td, f) for f in ("mesh.dat", "inlet.dat", "wall.dat")]


# the writer's / loader's arrays and the reference file's dataset names
STRUCTURE = [("node_pos", "n_pos"), ("c2n", "c_v_n_map"), ("c2c", "c_v_c_map"), ("node_type", "n_type"),
             ("if2c", "if_v_c_map"), ("if2n", "if_v_n_map"), ("if_area", "iface_area"), ("if_normal", "iface_norm"),
             ("if_u", "iface_u_norm"), ("if_v", "iface_v_norm")]
GEOMETRY = [("c_det", "c_det"), ("c_vol", "c_volume"), ("node_vol", "n_vol")]


@pytest.mark.parametrize("name", NAMES)
def test_box12_writer_reproduces_reference_mesh_structure(name):
    """Nodes, tets, face neighbours, node types and inlet faces of the reference's shipped mesh, bit for
    bit, from the 12-tet box writer and the text-mesh loader."""
    ds = MESHES[name]["datasets"]
    with tempfile.TemporaryDirectory() as td:
        m = oppx.mesh_from_files(*box_files(name, td))
    for ours, ref in STRUCTURE:
        a = np.ascontiguousarray(m[ours]).astype(np.dtype(ds[ref]["dtype"])).reshape(ds[ref]["shape"])
        assert sha(a) == ds[ref]["sha256"], f"{name}: {ours} vs {ref}"


@pytest.fixture(scope="module")
def ref_files():
    if not os.path.isdir(TARS):
        pytest.skip("the reference tree (its HDF5 meshes) is not here")
    with tempfile.TemporaryDirectory() as td:
        paths = {}
        for name in NAMES:
            with tarfile.open(os.path.join(TARS, name + ".tar.gz")) as t:
                data = t.extractfile(name + ".hdf5").read()
            paths[name] = os.path.join(td, name + ".hdf5")
            with open(paths[name], "wb") as fh:
                fh.write(data)
        yield paths


@pytest.mark.parametrize("name", NAMES)
def test_reader_matches_independent_reader_on_reference_files(ref_files, name):
    """Every dataset of the reference's file through the library's reader: dtype, shape and elements
    as the pure-Python reader read them; the reference's set-size and map/dat extents as
    get_dataset_properties takes them (rows = first extent, dim = second or 1)."""
    path = ref_files[name]
    for k, d in MESHES[name]["datasets"].items():
        a = oppx.h5_dataset(path, k)
        assert a is not None, k
        shape = d["shape"] if len(d["shape"]) == 2 else [d["shape"][0] if d["shape"] else 1, 1]
        assert list(a.shape) == shape and a.dtype == np.dtype(d["dtype"]), k
        if "sha256" in d:
            assert sha(a) == d["sha256"], k
    assert oppx.h5_dataset(path, "part_position") is None  # the particle dats the app declares are absent
    assert int(oppx.h5_dataset(path, "mesh_cells")[0, 0]) == MESHES[name]["datasets"]["c_det"]["shape"][0]


@pytest.mark.parametrize("name", NAMES)
def test_reference_geometry_is_the_reference_loader_without_fma(ref_files, name):
    """c_det, c_volume and n_vol of the shipped file: the reference loader's arithmetic without FMA
    contraction, bit for bit (oracle/_ref/mesh_ref_nofma on the writer's text files); the library's
    loader, which follows the reference's contracting build, within 1e-12 relative (measured: 4e-14 on
    c_det, 1.6e-13 on the volumes, whose determinants cancel)."""
    if not os.path.exists(MESH_REF_NOFMA):
        pytest.skip("oracle/_ref/mesh_ref_nofma not built (needs the reference tree)")
    path = ref_files[name]
    ds = MESHES[name]["datasets"]
    with tempfile.TemporaryDirectory() as td:
        files = box_files(name, td)
        out = os.path.join(td, "out")
        os.makedirs(out)
        subprocess.run([MESH_REF_NOFMA, *files, out], check=True, capture_output=True, timeout=300)
        m = oppx.mesh_from_files(*files)
        for ours, ref in GEOMETRY:
            nofma = np.fromfile(os.path.join(out, ours + ".bin"), np.float64).reshape(ds[ref]["shape"])
            assert sha(nofma) == ds[ref]["sha256"], f"{name}: {ours} vs {ref}"
            stored = oppx.h5_dataset(path, ref).reshape(ds[ref]["shape"])
            lib = np.asarray(m[ours], np.float64).reshape(ds[ref]["shape"])
            scale = np.abs(stored).max()
            assert np.abs(lib - stored).max() <= 1e-12 * scale, f"{name}: {ours}"


def test_reader_conversions_and_errors():
    rng = np.random.default_rng(3)
    ds = {"ints": rng.integers(-1000, 1000, (7, 3)).astype(np.int32), "wide": np.array([[1], [-5], [2 ** 40]], np.int64),
          "reals": rng.standard_normal((4, 2)) * 100, "single": rng.standard_normal((5,)).astype(np.float32),
          "count": np.array([7], np.int32), "with space": np.arange(3, dtype=np.float64)}
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "t.h5")
        h5write.write(p, ds)
        for k, v in ds.items():
            assert np.array_equal(oppx.h5_dataset(p, k).ravel(), v.ravel()), k
        assert np.array_equal(oppx.h5_dataset(p, "ints", np.float64), ds["ints"].astype(np.float64))
        assert np.array_equal(oppx.h5_dataset(p, "reals", np.int32), np.trunc(ds["reals"]).astype(np.int32))
        assert np.array_equal(oppx.h5_dataset(p, "single", np.float64).ravel(), ds["single"].astype(np.float64))
        assert oppx.h5_dataset(p, "missing") is None
        with pytest.raises(oppx.OppxError, match="int32 range"):
            oppx.h5_dataset(p, "wide", np.int32)
        # fewer rows than stored is allowed, a different dim is not
        out = np.zeros((2, 3), np.int32)
        lib = oppx.lib()
        assert lib.oppx_h5_read(p.encode(), b"ints", oppx.H5_CODES["int32"], 2, 3, out.ctypes.data) == 0
        assert np.array_equal(out, ds["ints"][:2])
        assert lib.oppx_h5_read(p.encode(), b"ints", oppx.H5_CODES["int32"], 2, 4, out.ctypes.data) == 1
        assert lib.oppx_h5_read(p.encode(), b"ints", oppx.H5_CODES["int32"], 8, 3, out.ctypes.data) == 1
        # not an HDF5 file, a truncated one, a missing one: errors with a reason, never a crash
        bad = os.path.join(td, "bad.h5")
        with open(bad, "wb") as fh:
            fh.write(b"not hdf5 at all" * 10)
        with pytest.raises(oppx.OppxError, match="not an HDF5 file"):
            oppx.h5_dataset(bad, "ints")
        raw = open(p, "rb").read()
        for cut in (40, 200, len(raw) // 2, len(raw) - 100):
            with open(bad, "wb") as fh:
                fh.write(raw[:cut])
            try:
                oppx.h5_dataset(bad, "ints")
            except oppx.OppxError:
                pass
        with pytest.raises(oppx.OppxError, match="cannot open"):
            oppx.h5_dataset(os.path.join(td, "absent.h5"), "ints")
