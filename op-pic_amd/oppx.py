"""Python binding of liboppx (the C-ABI in include/oppx.h) over ctypes.

Host builders return numpy arrays in the reference host layout (AoS). Device state lives in
torch tensors (torch is used only for device memory, streams and events); every compute call
goes to a gfx950 kernel in liboppx.so, and a missing library raises instead of falling back.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_double, c_int, c_int32, c_int64, c_size_t, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboppx.so")

MAX_CELL_INDEX = 2147483647
OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_STATE = 0, 1, 2, 3, 4


class TetAux(Structure):
    _fields_ = [("c2", c_double), ("c2c", c_int32 * 4), ("pad", c_int32 * 2)]


class TetMesh(Structure):
    _fields_ = [("n_cells", c_int64), ("n_nodes", c_int64), ("det", c_void_p), ("aux", c_void_p),
                ("c2n", c_void_p), ("ef", c_void_p)]


class DhGrid(Structure):
    _fields_ = [("cell", c_void_p), ("dims", c_int64 * 3), ("min", c_double * 3), ("inv_gs", c_double)]


class FempicParts(Structure):
    _fields_ = [("stride", c_int64), ("pos", c_void_p), ("vel", c_void_p), ("lc", c_void_p), ("p2c", c_void_p)]


class MoveCounters(Structure):
    _fields_ = [("size", c_int32), ("removed", c_int32), ("holes_low", c_int32), ("max_hops", c_int32),
                ("total_hops", c_int64), ("faults", c_int32), ("pad", c_int32), ("queue", c_int32 * 8)]


COUNTER_WORDS = 16  # int32 words of one oppx_move_counters


class QuadCell(Structure):
    _fields_ = [("llx", c_double), ("lly", c_double), ("c2c", c_int32 * 4)]


class NesoParts(Structure):
    _fields_ = [("stride", c_int64), ("pos", c_void_p), ("vel", c_void_p), ("mdir", c_void_p), ("p2c", c_void_p)]


class Tuning(Structure):
    _fields_ = [("tile", c_int32), ("refill_min", c_int32), ("split", c_int32), ("persistent", c_int32),
                ("sort_items", c_int32), ("inline_hops", c_int32), ("first_block", c_int32), ("lanes", c_int32)]


class BatchInst(Structure):
    _fields_ = [("cell_base", c_int64), ("n_cells", c_int64), ("node_base", c_int64), ("n_nodes", c_int64),
                ("dh_cell", c_void_p), ("dims", c_int64 * 3), ("min", c_double * 3), ("inv_gs", c_double),
                ("k", c_double), ("dt", c_double)]


class Instance(Structure):
    """oppx_instance: one FEM-PIC instance as host arrays (the oppx_arena_admit input)."""
    _fields_ = [("n_cells", c_int64), ("n_nodes", c_int64), ("n_part", c_int64), ("c_det", c_void_p),
                ("c_vol", c_void_p), ("c2c", c_void_p), ("c2n", c_void_p), ("c_ef", c_void_p), ("pos", c_void_p),
                ("vel", c_void_p), ("p2c", c_void_p), ("overlay", c_void_p), ("dh_dims", c_int64 * 3),
                ("dh_min", c_double * 3), ("dh_inv_gs", c_double), ("charge", c_double), ("mass", c_double),
                ("dt", c_double)]


class Inlet(Structure):
    _fields_ = [("n_faces", c_int64), ("face_cell", c_void_p), ("u", c_void_p), ("v", c_void_p), ("normal", c_void_p),
                ("node_pos", c_void_p)]


class CabanaParts(Structure):
    _fields_ = [("stride", c_int64), ("pos", c_void_p), ("vel", c_void_p), ("streak", c_void_p), ("weight", c_void_p),
                ("p2c", c_void_p)]


class CabanaMesh(Structure):
    _fields_ = [("n_cells", c_int64), ("c2c6", c_void_p), ("interp", c_void_p)]


class DatDesc(Structure):
    _fields_ = [("base", c_void_p), ("dim", c_int32), ("elem_bytes", c_int32)]


assert ctypes.sizeof(TetAux) == 32 and ctypes.sizeof(MoveCounters) == 4 * COUNTER_WORDS
assert ctypes.sizeof(BatchInst) == 112 and ctypes.sizeof(Tuning) == 32

# name -> (restype, argtypes): every symbol include/oppx.h declares
_SIGS = {
    "oppx_init": (c_int, [c_int]),
    "oppx_last_error": (ctypes.c_char_p, []),
    "oppx_device_malloc": (c_int, [POINTER(c_void_p), c_size_t]),
    "oppx_device_free": (c_int, [c_void_p]),
    "oppx_memcpy_htod": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "oppx_memcpy_dtoh": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "oppx_memcpy_dtod": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "oppx_memset_zero": (c_int, [c_void_p, c_size_t, c_void_p]),
    "oppx_stream_sync": (c_int, [c_void_p]),
    "oppx_offset_indices": (c_int, [c_void_p, c_int64, c_int, c_int, c_int, c_int32, c_void_p]),
    "oppx_fill_pattern": (c_int, [c_void_p, c_int64, c_int, c_int, c_void_p, c_int64, c_void_p]),
    "oppx_aos_to_soa": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_int64, c_int, c_void_p]),
    "oppx_soa_to_aos": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_int64, c_int, c_void_p]),
    "oppx_box_mesh_sizes": (c_int, [c_int, c_int, c_int, POINTER(c_int64), POINTER(c_int64)]),
    "oppx_box_mesh_build": (c_int, [c_int, c_int, c_int, c_double, c_double, c_double, c_int,
                                    c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "oppx_box_field_sines": (c_int, [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "oppx_dh_overlay_filename": (c_int, [c_int64, c_double, c_void_p, ctypes.c_char_p, ctypes.c_size_t]),
    "oppx_dh_overlay_save": (c_int, [ctypes.c_char_p, c_void_p, c_int64]),
    "oppx_dh_overlay_load": (c_int, [ctypes.c_char_p, c_void_p, c_int64]),
    "oppx_box_particles_uniform": (c_int, [c_int64, c_uint64, c_double, c_int, c_int, c_int,
                                           c_double, c_double, c_double, c_int, c_void_p, c_void_p, c_void_p]),
    "oppx_dh_grid_dims": (c_int, [c_void_p, c_void_p, c_double, c_void_p]),
    "oppx_dh_build_overlay_nodes": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                            c_void_p, c_void_p, c_double, c_void_p]),
    "oppx_neso_mesh": (c_int, [c_int, c_int, c_double, c_void_p, c_void_p]),
    "oppx_neso_particles": (c_int, [c_int64, c_uint64, c_int, c_int, c_double, c_double, c_void_p, c_void_p,
                                    c_void_p]),
    "oppx_dh_build_overlay_geom": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                                           c_void_p]),
    "oppx_tet_aux_build": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "oppx_fempic_push": (c_int, [c_int64, POINTER(FempicParts), POINTER(TetMesh), c_double, c_double,
                                 c_void_p, c_void_p]),
    "oppx_fempic_work_bytes": (ctypes.c_size_t, [c_int64]),
    "oppx_fempic_move": (c_int, [c_int64, POINTER(FempicParts), POINTER(TetMesh), POINTER(DhGrid),
                                 c_void_p, c_void_p, c_void_p, ctypes.c_size_t, c_void_p, c_void_p]),
    "oppx_fempic_deposit_atomic": (c_int, [c_int64, POINTER(FempicParts), POINTER(TetMesh), c_void_p,
                                           c_void_p, c_void_p]),
    "oppx_fempic_step_fused": (c_int, [c_int64, POINTER(FempicParts), POINTER(TetMesh), POINTER(DhGrid),
                                       c_double, c_double, c_void_p, c_void_p, c_void_p, ctypes.c_size_t, c_void_p,
                                       c_void_p, c_void_p]),
    "oppx_fempic_step_fused_gather": (c_int, [c_int64, c_void_p, POINTER(FempicParts), POINTER(FempicParts),
                                              POINTER(TetMesh), POINTER(DhGrid), c_double, c_double, c_void_p,
                                              c_void_p, c_void_p, ctypes.c_size_t, c_void_p, c_void_p, c_void_p]),
    "oppx_gather_dats": (c_int, [c_int64, c_int64, c_void_p, c_void_p, POINTER(DatDesc), POINTER(DatDesc), c_int,
                                 c_void_p]),
    "oppx_fempic_step_fused_rec": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, POINTER(FempicParts), c_void_p,
                                           POINTER(TetMesh), POINTER(DhGrid), c_double, c_double, c_void_p, c_void_p,
                                           c_void_p, c_size_t, c_void_p, c_void_p, c_void_p]),
    "oppx_gather_records": (c_int, [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "oppx_fempic_batch_step_fused": (c_int, [c_int64, POINTER(FempicParts), POINTER(TetMesh), c_void_p, c_void_p,
                                             c_int32, c_void_p, c_void_p, c_void_p, ctypes.c_size_t, c_void_p,
                                             c_void_p, c_void_p]),
    "oppx_fempic_batch_step_fused_rec": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, POINTER(FempicParts),
                                                 c_void_p, POINTER(TetMesh), c_void_p, c_void_p, c_int32, c_void_p,
                                                 c_void_p, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p]),
    "oppx_quad_cells_build": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "oppx_neso_step": (c_int, [c_int64, POINTER(NesoParts), c_void_p, c_int64, c_double, c_double, c_double, c_double,
                               c_int, c_void_p, c_void_p, c_void_p]),
    "oppx_quad_grid_check": (c_int, [c_void_p, c_int64, c_int, c_int, c_double, c_void_p, c_void_p]),
    "oppx_neso_verify": (c_int, [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_int, c_double, c_void_p,
                                 c_void_p, c_void_p]),
    "oppx_fempic_inject": (c_int, [c_int64, c_int64, POINTER(FempicParts), POINTER(Inlet), c_int64, c_void_p, c_void_p,
                                   c_int64, c_double, c_double, c_double, c_double, c_void_p, c_void_p]),
    "oppx_fempic_node_charge_density": (c_int, [c_int64, c_void_p, c_void_p, c_double, c_void_p]),
    "oppx_fempic_electric_field": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                                           c_void_p]),
    "oppx_fempic_boundary_potential": (c_int, [c_int64, c_void_p, c_double, c_void_p, c_void_p]),
    "oppx_max_scratch_bytes": (c_size_t, [c_int64]),
    "oppx_cabana_move_deposit": (c_int, [c_int64, POINTER(CabanaParts), POINTER(CabanaMesh), c_void_p, c_double,
                                         c_double, POINTER(c_double), c_void_p, c_void_p, c_void_p]),
    "oppx_max_reduce": (c_int, [c_int64, c_void_p, c_int, c_double, c_void_p, c_void_p, c_size_t, c_void_p]),
    "oppx_neso_step_grid": (c_int, [c_int64, POINTER(NesoParts), c_int, c_int, c_double, c_double, c_double, c_double,
                                    c_int, c_void_p, c_void_p, c_void_p]),
    "oppx_fempic_step_begin": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "oppx_hole_bits_words": (c_int64, [c_int64]),
    "oppx_node_cell_csr": (c_int, [c_int64, c_int64, c_void_p, c_void_p, c_void_p]),
    "oppx_fempic_deposit_scratch_bytes": (ctypes.c_size_t, [c_int64]),
    "oppx_fempic_deposit_sorted_f32": (c_int, [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                                               c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                               ctypes.c_size_t, c_void_p]),
    "oppx_fempic_deposit_sorted_f32_perm": (c_int, [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                                    c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64,
                                                    c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_size_t,
                                                    c_void_p]),
    "oppx_lower_bound": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "oppx_sort_scratch_bytes": (c_size_t, [c_int64]),
    "oppx_sort_by_cell": (c_int, [c_int64, c_int64, c_void_p, c_int64, c_void_p, POINTER(DatDesc), POINTER(DatDesc),
                                  c_int, c_void_p, c_int, c_void_p, c_size_t, c_void_p]),
    "oppx_set_tuning": (c_int, [POINTER(Tuning)]),
    "oppx_get_tuning": (c_int, [POINTER(Tuning)]),
    "oppx_box12_write": (c_int, [c_int, c_int, c_int, c_double, ctypes.c_char_p]),
    "oppx_box12_particles": (c_int, [c_int64, ctypes.c_uint64, c_double, c_int, c_int, c_int, c_double, c_double,
                                     c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "oppx_mesh_file_sizes": (c_int, [ctypes.c_char_p, ctypes.c_char_p, c_void_p, c_void_p, c_void_p]),
    "oppx_mesh_file_load": (c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p] + [c_void_p] * 13),
    "oppx_h5_dataset_info": (c_int, [ctypes.c_char_p, ctypes.c_char_p, c_void_p, c_void_p, c_void_p]),
    "oppx_h5_read": (c_int, [ctypes.c_char_p, ctypes.c_char_p, c_int32, c_int64, c_int64, c_void_p]),
    "oppx_holefill_scratch_bytes": (c_size_t, [c_int64]),
    "oppx_holefill": (c_int, [c_int64, c_int64, c_void_p, c_void_p, POINTER(DatDesc), c_int, c_void_p,
                              c_size_t, c_void_p]),
    "oppx_holefill_noalloc": (c_int, [c_int64, c_int64, c_void_p, c_void_p, POINTER(DatDesc), c_int, c_void_p]),
    "oppx_arena_create": (c_int, [c_int64, c_int64, c_int64, c_int64, c_int32, c_void_p, POINTER(c_void_p)]),
    "oppx_arena_destroy": (c_int, [c_void_p]),
    "oppx_arena_admit": (c_int, [c_void_p, POINTER(Instance), POINTER(c_int32)]),
    "oppx_arena_step": (c_int, [c_void_p]),
    "oppx_arena_retire": (c_int, [c_void_p, c_int32, c_int64, POINTER(c_int64), c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p]),
    "oppx_arena_counters": (c_int, [c_void_p, POINTER(MoveCounters)]),
}

_lib = None


def lib():
    """Load liboppx.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        # One HIP runtime per process: torch bundles libamdhip64 (SONAME libamdhip64.so.7). Loading
        # torch first makes liboppx.so's libamdhip64.so.7 dependency bind to that same copy.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return list(_SIGS)


class OppxError(RuntimeError):
    pass


def check(rc: int, what: str = ""):
    if rc != OK:
        msg = lib().oppx_last_error().decode(errors="replace")
        raise OppxError(f"{what} failed (rc={rc}): {msg}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(c_void_p)


# ---------------------------------------------------------------------------------------------
# host builders
# ---------------------------------------------------------------------------------------------

ORDERS = {"x": 0, "morton": 1}


def box_mesh(nx: int, ny: int, nz: int, lx=1.0, ly=1.0, lz=1.0, order="x") -> dict:
    """Box of nx*ny*nz cubes x 6 tets with the FEM geometry the mover consumes; cubes and nodes
    numbered x fastest ("x") or in Morton order ("morton")."""
    L = lib()
    nn, nc = c_int64(), c_int64()
    check(L.oppx_box_mesh_sizes(nx, ny, nz, ctypes.byref(nn), ctypes.byref(nc)), "box_mesh_sizes")
    nn, nc = nn.value, nc.value
    m = dict(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz, order=order,
             node_pos=np.zeros((nn, 3)), c2n=np.zeros((nc, 4), np.int32), c2c=np.zeros((nc, 4), np.int32),
             c_det=np.zeros((nc, 16)), c_vol=np.zeros(nc))
    check(L.oppx_box_mesh_build(nx, ny, nz, lx, ly, lz, ORDERS[order], _ptr(m["node_pos"]), _ptr(m["c2n"]),
                                _ptr(m["c2c"]),
                                _ptr(m["c_det"]), _ptr(m["c_vol"])), "box_mesh_build")
    m["n_nodes"], m["n_cells"] = nn, nc
    return m


def box12_write(nx: int, ny: int, nz: int, cube_length: float, directory: str):
    """mesh.dat / inlet.dat / wall.dat of the 12-tet box mesh (cube_length in mm) in `directory`."""
    check(lib().oppx_box12_write(nx, ny, nz, cube_length, directory.encode()), "box12_write")


def box12_mesh(nx: int, ny: int, nz: int, lx=1.0, ly=1.0, lz=1.0) -> dict:
    """The 12-tet box mesh of nx x ny x nz cubes on [0,lx]x[0,ly]x[0,lz] (m), written in the app's text
    format (cube edges in mm; a cubic cell needs lx/nx == ly/ny == lz/nz) and read back with
    mesh_from_files: the geometry the reference app's loader computes from those files."""
    import tempfile
    cube_mm = 1000.0 * lx / nx
    if abs(1000.0 * ly / ny - cube_mm) > 1e-12 * cube_mm or abs(1000.0 * lz / nz - cube_mm) > 1e-12 * cube_mm:
        raise ValueError("the 12-tet box mesh has cubic cells: lx/nx, ly/ny and lz/nz must agree")
    with tempfile.TemporaryDirectory() as td:
        box12_write(nx, ny, nz, cube_mm, td)
        m = mesh_from_files(os.path.join(td, "mesh.dat"), os.path.join(td, "inlet.dat"), os.path.join(td, "wall.dat"))
    m.update(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz, split=12)
    return m


def box12_particles(n: int, seed: int, sigma: float, m: dict):
    pos, vel = np.zeros((n, 3)), np.zeros((n, 3))
    p2c = np.zeros(n, np.int32)
    check(lib().oppx_box12_particles(n, seed, sigma, m["nx"], m["ny"], m["nz"], m["lx"], m["ly"], m["lz"],
                                     _ptr(m["c_det"]), _ptr(m["c_vol"]), _ptr(pos), _ptr(vel), _ptr(p2c)),
          "box12_particles")
    return pos, vel, p2c


H5_TYPES = {1: np.int32, 2: np.int64, 3: np.float32, 4: np.float64}  # OPPX_H5_* element types
H5_CODES = {np.dtype(v).name: k for k, v in H5_TYPES.items()}


def h5_dataset(path: str, name: str, dtype=None):
    """One dataset of an OP-PIC HDF5 file (oppx_h5_read) as a (rows, dim) array, in its stored element
    type or converted to dtype (int32, int64, float32 or float64). None when the file has no such
    dataset."""
    L = lib()
    t, rows, dim = c_int32(), c_int64(), c_int64()
    rc = L.oppx_h5_dataset_info(path.encode(), name.encode(), ctypes.byref(t), ctypes.byref(rows), ctypes.byref(dim))
    if rc == 5:  # OPPX_ERR_NOTFOUND
        return None
    check(rc, "h5_dataset_info")
    want = np.dtype(dtype) if dtype is not None else np.dtype(H5_TYPES[t.value])
    code = {np.dtype(v): k for k, v in H5_TYPES.items()}[want]
    out = np.zeros((rows.value, dim.value), want)
    check(L.oppx_h5_read(path.encode(), name.encode(), code, rows.value, dim.value, _ptr(out)), "h5_read")
    return out


def mesh_from_files(mesh_path: str, inlet_path: str | None = None, wall_path: str | None = None) -> dict:
    """A FEM-PIC mesh read from the app's text files (the box_mesh dict layout, plus node volumes and
    types and the inlet faces)."""
    enc = lambda p: p.encode() if p else None
    sizes = [c_int64(), c_int64(), c_int64()]
    L = lib()
    check(L.oppx_mesh_file_sizes(enc(mesh_path), enc(inlet_path), *[ctypes.byref(v) for v in sizes]), "mesh_file_sizes")
    nn, nc, nf = (int(v.value) for v in sizes)
    m = dict(n_nodes=nn, n_cells=nc, n_ifaces=nf, node_pos=np.zeros((nn, 3)), node_type=np.zeros(nn, np.int32),
             node_vol=np.zeros(nn), c2n=np.zeros((nc, 4), np.int32), c2c=np.zeros((nc, 4), np.int32),
             c_det=np.zeros((nc, 16)), c_vol=np.zeros(nc), if2c=np.zeros(max(nf, 1), np.int32),
             if2n=np.zeros((max(nf, 1), 3), np.int32), if_u=np.zeros((max(nf, 1), 3)), if_v=np.zeros((max(nf, 1), 3)),
             if_normal=np.zeros((max(nf, 1), 3)), if_area=np.zeros(max(nf, 1)))
    keys = ["node_pos", "node_type", "node_vol", "c2n", "c2c", "c_det", "c_vol", "if2c", "if2n", "if_u", "if_v",
            "if_normal", "if_area"]
    check(L.oppx_mesh_file_load(enc(mesh_path), enc(inlet_path), enc(wall_path), *[_ptr(m[k]) for k in keys]),
          "mesh_file_load")
    for k in keys[7:]:
        m[k] = m[k][:nf]
    return m


def box_field_sines(m: dict) -> np.ndarray:
    ef = np.zeros((m["n_cells"], 3))
    check(lib().oppx_box_field_sines(m["n_nodes"], m["n_cells"], _ptr(m["node_pos"]), _ptr(m["c2n"]),
                                     _ptr(m["c_det"]), _ptr(m["c_vol"]), _ptr(ef)), "box_field_sines")
    return ef


def box_particles(n: int, seed: int, sigma: float, m: dict):
    pos, vel = np.zeros((n, 3)), np.zeros((n, 3))
    p2c = np.zeros(n, np.int32)
    check(lib().oppx_box_particles_uniform(n, seed, sigma, m["nx"], m["ny"], m["nz"], m["lx"], m["ly"], m["lz"],
                                           ORDERS[m.get("order", "x")], _ptr(pos), _ptr(vel), _ptr(p2c)),
          "box_particles_uniform")
    return pos, vel, p2c


def dh_dims(bmin, bmax, gs):
    bmin = np.ascontiguousarray(bmin, np.float64)
    bmax = np.ascontiguousarray(bmax, np.float64)
    dims = np.zeros(3, np.int64)
    check(lib().oppx_dh_grid_dims(_ptr(bmin), _ptr(bmax), gs, _ptr(dims)), "dh_grid_dims")
    return dims


def dh_overlay(m: dict, bmin, bmax, gs) -> np.ndarray:
    """Direct-hop overlay (owner tet per structured box, flattened x-fastest)."""
    bmin = np.ascontiguousarray(bmin, np.float64)
    bmax = np.ascontiguousarray(bmax, np.float64)
    dims = dh_dims(bmin, bmax, gs)
    ov = np.zeros(int(np.prod(dims)), np.int32)
    check(lib().oppx_dh_build_overlay_nodes(m["n_cells"], _ptr(m["c_det"]), _ptr(m["c_vol"]), None,
                                            _ptr(m["c2n"]), _ptr(m["node_pos"]), _ptr(bmin), _ptr(bmax), gs,
                                            _ptr(ov)), "dh_build_overlay")
    return ov


def neso_mesh(nx: int, ny: int, cell_width: float) -> dict:
    """2-D periodic quad mesh of the NESO-advection app."""
    nc = nx * ny
    m = dict(nx=nx, ny=ny, cell_width=cell_width, n_cells=nc, c_pos_ll=np.zeros((nc, 2)),
             c2c=np.zeros((nc, 4), np.int32))
    check(lib().oppx_neso_mesh(nx, ny, cell_width, _ptr(m["c_pos_ll"]), _ptr(m["c2c"])), "neso_mesh")
    return m


def neso_particles(n: int, seed: int, m: dict, vmax: float):
    pos, vel = np.zeros((n, 2)), np.zeros((n, 2))
    p2c = np.zeros(n, np.int32)
    check(lib().oppx_neso_particles(n, seed, m["nx"], m["ny"], m["cell_width"], vmax, _ptr(pos), _ptr(vel),
                                    _ptr(p2c)), "neso_particles")
    return pos, vel, p2c


TUNING_FIELDS = ("tile", "refill_min", "split", "persistent", "sort_items", "inline_hops", "first_block", "lanes")


def set_tuning(tile=256, refill_min=16, split=0, persistent=0, sort_items=16, inline_hops=64, first_block=256,
               lanes=1):
    """Process-wide launch shape of the walk and the sort (results do not depend on it)."""
    t = Tuning(tile, refill_min, split, persistent, sort_items, inline_hops, first_block, lanes)
    check(lib().oppx_set_tuning(ctypes.byref(t)), "set_tuning")


def get_tuning() -> dict:
    t = Tuning()
    check(lib().oppx_get_tuning(ctypes.byref(t)), "get_tuning")
    return {f: int(getattr(t, f)) for f in TUNING_FIELDS}


def node_cell_csr(m: dict):
    start = np.zeros(m["n_nodes"] + 1, np.int32)
    ref = np.zeros(4 * m["n_cells"], np.int32)
    check(lib().oppx_node_cell_csr(m["n_cells"], m["n_nodes"], _ptr(np.ascontiguousarray(m["c2n"], np.int32)),
                                   _ptr(start), _ptr(ref)), "node_cell_csr")
    return start, ref


def dh_overlay_geom(m: dict, bmin, bmax, gs) -> np.ndarray:
    """Direct-hop overlay built from c_det / c_volume only."""
    bmin = np.ascontiguousarray(bmin, np.float64)
    bmax = np.ascontiguousarray(bmax, np.float64)
    ov = np.zeros(int(np.prod(dh_dims(bmin, bmax, gs))), np.int32)
    check(lib().oppx_dh_build_overlay_geom(m["n_cells"], _ptr(m["c_det"]), _ptr(m["c_vol"]), None, _ptr(bmin),
                                           _ptr(bmax), gs, _ptr(ov)), "dh_build_overlay_geom")
    return ov


def dh_overlay_filename(n_cells: int, gs: float, expansion=(0.0, 0.0, 0.0)) -> str:
    buf = ctypes.create_string_buffer(256)
    ex = np.ascontiguousarray(expansion, np.float64)
    check(lib().oppx_dh_overlay_filename(n_cells, gs, _ptr(ex), buf, 256), "dh_overlay_filename")
    return buf.value.decode()


def dh_overlay_save(path: str, overlay: np.ndarray):
    ov = np.ascontiguousarray(overlay, np.int32)
    check(lib().oppx_dh_overlay_save(path.encode(), _ptr(ov), ov.size), "dh_overlay_save")


def dh_overlay_load(path: str, n: int) -> np.ndarray:
    ov = np.zeros(n, np.int32)
    check(lib().oppx_dh_overlay_load(path.encode(), _ptr(ov), n), "dh_overlay_load")
    return ov


def sort_particles_by_cell(pos, vel, p2c):
    """Stable reorder of an input population by containing cell (an input-layout choice)."""
    order = np.argsort(p2c, kind="stable")
    return pos[order], vel[order], p2c[order]


# ---------------------------------------------------------------------------------------------
# device state of one FEM-PIC instance
# ---------------------------------------------------------------------------------------------

class FempicDevice:
    """Mesh + particles of one FEM-PIC instance resident on the GPU, driven through the C-ABI."""

    def __init__(self, mesh: dict, c_ef: np.ndarray, pos: np.ndarray, vel: np.ndarray, p2c: np.ndarray,
                 capacity: int | None = None, overlay: np.ndarray | None = None, bmin=None, gs: float | None = None,
                 bmax=None, device: str = "cuda"):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        L = lib()
        check(L.oppx_init(self.dev.index or 0), "oppx_init")
        n = int(pos.shape[0])
        cap = int(capacity) if capacity is not None else n
        if cap < n:
            raise ValueError("capacity below particle count")
        self.n0, self.cap = n, cap
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(self.dev, dtype=dt)
        nc, nn = mesh["n_cells"], mesh["n_nodes"]
        # host-side validation: every index a kernel dereferences must be in range
        if not ((mesh["c2n"] >= 0).all() and (mesh["c2n"] < nn).all()):
            raise ValueError("c2n entries out of [0, n_nodes)")
        if not ((mesh["c2c"] >= -1).all() and (mesh["c2c"] < nc).all()):
            raise ValueError("c2c entries out of [-1, n_cells)")
        if n and not ((p2c >= 0).all() and (p2c < nc).all()):
            raise ValueError("initial p2c entries out of [0, n_cells)")
        self.nc, self.nn = nc, nn
        self._mesh_host = dict(n_cells=nc, n_nodes=nn, c2n=mesh["c2n"])
        self.det = t(mesh["c_det"].reshape(-1), torch.float64)
        self.c_vol = t(mesh["c_vol"], torch.float64)
        self.c2c = t(mesh["c2c"].reshape(-1), torch.int32)
        self.c2n = t(mesh["c2n"].reshape(-1), torch.int32)
        self.ef = t(c_ef.reshape(-1), torch.float64)
        self.aux = torch.zeros(nc * 4, dtype=torch.float64, device=self.dev)  # 32 B per cell
        check(L.oppx_tet_aux_build(self.c_vol.data_ptr(), self.c2c.data_ptr(), nc, self.aux.data_ptr(),
                                   self.stream()), "tet_aux_build")
        self.mesh = TetMesh(nc, nn, self.det.data_ptr(), self.aux.data_ptr(), self.c2n.data_ptr(),
                            self.ef.data_ptr())
        # particles: SoA with stride = capacity
        self.pos = torch.zeros((3, cap), dtype=torch.float64, device=self.dev)
        self.vel = torch.zeros((3, cap), dtype=torch.float64, device=self.dev)
        self.lc = torch.zeros((4, cap), dtype=torch.float64, device=self.dev)
        self.p2c = torch.full((cap,), MAX_CELL_INDEX, dtype=torch.int32, device=self.dev)
        self.pos[:, :n] = t(pos.T, torch.float64)
        self.vel[:, :n] = t(vel.T, torch.float64)
        self.p2c[:n] = t(p2c, torch.int32)
        self.parts = FempicParts(cap, self.pos.data_ptr(), self.vel.data_ptr(), self.lc.data_ptr(),
                                 self.p2c.data_ptr())
        self.counters = torch.zeros(COUNTER_WORDS, dtype=torch.int32, device=self.dev)
        self.counters[0] = n
        self.hole_bits = torch.zeros(int(L.oppx_hole_bits_words(cap)), dtype=torch.int64, device=self.dev)
        self.work_bytes = int(L.oppx_fempic_work_bytes(cap))
        self.work = torch.zeros(max(self.work_bytes, 8), dtype=torch.uint8, device=self.dev)
        self.scratch_bytes = int(L.oppx_holefill_scratch_bytes(cap))
        self.scratch = torch.zeros(self.scratch_bytes, dtype=torch.uint8, device=self.dev)
        self.charge = torch.zeros(nn, dtype=torch.float64, device=self.dev)
        self.hops = torch.zeros(cap, dtype=torch.int32, device=self.dev)
        self.dh = None
        if overlay is not None:
            dims = dh_dims(bmin, bmax, gs)
            if overlay.size != int(np.prod(dims)):
                raise ValueError("overlay size does not match the grid dims")
            if not (((overlay >= 0) & (overlay < nc)) | (overlay == MAX_CELL_INDEX)).all():
                raise ValueError("overlay entries must be cells or MAX_CELL_INDEX")
            self.overlay = t(overlay, torch.int32)
            self.dh = DhGrid(self.overlay.data_ptr(), (c_int64 * 3)(*dims), (c_double * 3)(*bmin), 1.0 / gs)
        self.dats = (DatDesc * 4)(DatDesc(self.pos.data_ptr(), 3, 8), DatDesc(self.vel.data_ptr(), 3, 8),
                                  DatDesc(self.lc.data_ptr(), 4, 8), DatDesc(self.p2c.data_ptr(), 1, 4))
        self.n_max = n  # upper bound of the live count (sizes only shrink without injection)
        self.pending = False  # a lazy sort's permutation waits for the next pass (or materialize)
        self.staged = False   # positions/velocities live in a 48-byte record buffer (pass_fused(stage=True))

    def stream(self):
        return self.torch.cuda.current_stream(self.dev).cuda_stream

    # attributes a clone gets afresh (the rest, the read-only mesh tensors and host metadata, it shares)
    _PER_INSTANCE = ("pos", "vel", "lc", "p2c", "counters", "hole_bits", "work", "scratch", "charge", "hops", "_swap",
                     "perm", "sort_scratch", "_rec", "_rec_cur", "cell_partials", "charge_f32", "dep_scratch", "parts",
                     "dats")

    def clone(self):
        """Another instance over this one's mesh -- its device tensors shared, read-only -- with a device
        copy of this particle state and its own counters, charge and work buffers (no host upload)."""
        torch = self.torch
        self.materialize()
        g = object.__new__(type(self))
        g.__dict__.update({k: v for k, v in self.__dict__.items() if k not in self._PER_INSTANCE})
        g.pos, g.vel, g.lc, g.p2c = (t.clone() for t in (self.pos, self.vel, self.lc, self.p2c))
        g.counters = self.counters.clone()
        for name in ("hole_bits", "work", "scratch", "charge", "hops"):
            setattr(g, name, torch.zeros_like(getattr(self, name)))
        if hasattr(self, "node_start"):  # the FP32 deposit's node rows are shared, its outputs are not
            for name in ("cell_partials", "charge_f32", "dep_scratch"):
                setattr(g, name, torch.zeros_like(getattr(self, name)))
        g.pending = g.staged = False
        g._rebind()
        return g

    def _dh_ref(self):
        return ctypes.byref(self.dh) if self.dh is not None else None

    # -- the reference loops, one C-ABI call each --------------------------------------------
    def push(self, charge, mass, dt):
        self.materialize()
        k = charge / mass * dt
        check(lib().oppx_fempic_push(self.n_max, ctypes.byref(self.parts), ctypes.byref(self.mesh), k, dt,
                                     self.counters.data_ptr(), self.stream()), "fempic_push")

    def move(self, record_hops=True):
        self.materialize()
        check(lib().oppx_fempic_move(self.n_max, ctypes.byref(self.parts), ctypes.byref(self.mesh), self._dh_ref(),
                                     self.counters.data_ptr(), self.hole_bits.data_ptr(), self.work.data_ptr(),
                                     self.work_bytes, self.hops.data_ptr() if record_hops else None, self.stream()),
              "fempic_move")

    def holefill(self):
        self.materialize()
        check(lib().oppx_holefill(self.n_max, self.cap, self.counters.data_ptr(), self.hole_bits.data_ptr(),
                                  self.dats, 4, self.scratch.data_ptr(), self.scratch_bytes, self.stream()),
              "holefill")

    def _swap_buffers(self):
        torch = self.torch
        if not hasattr(self, "_swap"):
            self._swap = [torch.empty_like(t) for t in (self.pos, self.vel, self.lc, self.p2c)]
            self.sort_scratch_bytes = int(lib().oppx_sort_scratch_bytes(self.cap))
            self.sort_scratch = torch.zeros(self.sort_scratch_bytes, dtype=torch.uint8, device=self.dev)
            self.perm = torch.zeros(self.cap, dtype=torch.int32, device=self.dev)
        return self._swap

    _DAT_SHAPES = {"pos": (3, 8), "vel": (3, 8), "lc": (4, 8), "p2c": (1, 4)}

    def _descs(self, tensors, names=("pos", "vel", "lc", "p2c")):
        shapes = [self._DAT_SHAPES[n] for n in names]
        assert len(tensors) == len(shapes)
        return (DatDesc * len(tensors))(*[DatDesc(t.data_ptr(), d, e) for t, (d, e) in zip(tensors, shapes)])

    def _records(self):
        """The two 48-byte position/velocity record buffers (6 doubles per particle slot)."""
        if not hasattr(self, "_rec"):
            self._rec = [self.torch.empty(6 * self.cap, dtype=self.torch.float64, device=self.dev) for _ in range(2)]
            self._rec_cur = 0
        return self._rec

    def sort(self, drop_removed=False, perm=None, lazy=False):
        """Stable sort of the particle dats by cell (removed last); drop_removed applies the finalize.
        lazy=True computes only the permutation (and the sorted p2c, written by the sort's last pass into
        the swap buffer) and leaves the rest of the record gather pending: the next fused step reads its
        inputs through it (pass_fused), or materialize() applies it. A lazy sort after a staged pass
        (pass_fused(stage=True)) keeps the position/velocity records: the next step reads them through
        the permutation."""
        if not (lazy and getattr(self, "staged", False) and not getattr(self, "pending", False)):
            self.materialize()
        swap = self._swap_buffers()
        cur = [self.pos, self.vel, self.lc, self.p2c]
        out_perm = perm if perm is not None else (self.perm if lazy else None)
        names = ("p2c",) if lazy else ("pos", "vel", "lc", "p2c")
        ins, outs = (cur[3:], swap[3:]) if lazy else (cur, swap)
        check(lib().oppx_sort_by_cell(self.n_max, self.cap, self.counters.data_ptr(), self.nc, self.p2c.data_ptr(),
                                      self._descs(ins, names), self._descs(outs, names), len(ins),
                                      out_perm.data_ptr() if out_perm is not None else None,
                                      1 if drop_removed else 0, self.sort_scratch.data_ptr(), self.sort_scratch_bytes,
                                      self.stream()), "sort_by_cell")
        if lazy:
            self.pending = True
            return
        self.pos, self.vel, self.lc, self.p2c = swap
        self._swap = cur
        self._rebind()

    def materialize(self):
        """Apply a pending permutation to the particle dats, and bring staged position/velocity records
        back into the SoA dats."""
        pending, staged = getattr(self, "pending", False), getattr(self, "staged", False)
        if staged and not pending:  # records of a staged pass, no sort since: same order, into pos/vel
            check(lib().oppx_gather_records(self.n_max, self.cap, self.counters.data_ptr(), None,
                                            self._rec[self._rec_cur].data_ptr(), self.pos.data_ptr(),
                                            self.vel.data_ptr(), self.stream()), "gather_records")
            self.staged = False
            return
        if not pending:
            return
        swap = self._swap_buffers()
        cur = [self.pos, self.vel, self.lc, self.p2c]
        # the sort already wrote the sorted p2c into the swap buffer
        if staged:
            check(lib().oppx_gather_records(self.n_max, self.cap, self.counters.data_ptr(), self.perm.data_ptr(),
                                            self._rec[self._rec_cur].data_ptr(), swap[0].data_ptr(),
                                            swap[1].data_ptr(), self.stream()), "gather_records")
            check(lib().oppx_gather_dats(self.n_max, self.cap, self.counters.data_ptr(), self.perm.data_ptr(),
                                         self._descs(cur[2:3], ("lc",)), self._descs(swap[2:3], ("lc",)), 1,
                                         self.stream()), "gather_dats")
        else:
            check(lib().oppx_gather_dats(self.n_max, self.cap, self.counters.data_ptr(), self.perm.data_ptr(),
                                         self._descs(cur[:3], ("pos", "vel", "lc")),
                                         self._descs(swap[:3], ("pos", "vel", "lc")), 3, self.stream()), "gather_dats")
        self.pos, self.vel, self.lc, self.p2c = swap
        self._swap = cur
        self._rebind()
        self.pending = False
        self.staged = False

    def deposit_sorted_f32(self):
        """FP32 Kahan deposit over the (cell-sorted) particles; returns the FP32 node charge tensor.
        Over a pending (lazy) sort the particle dats stay as they are: the sort left the sorted p2c in the
        swap buffer, and the deposit reads it with lc through the sort's permutation
        (oppx_fempic_deposit_sorted_f32_perm, the same sums bit for bit)."""
        torch = self.torch
        pending = getattr(self, "pending", False)
        if pending:
            swap = self._swap_buffers()
            # the swap lc buffer (32 B a particle, overwritten by the next pass) takes the sorted FP32 lc
            p2c, lc, lc_f32 = swap[3], self.lc, swap[2]
        else:
            self.materialize()
            p2c, lc = self.p2c, self.lc
        if not hasattr(self, "node_start"):
            st, ref = node_cell_csr(self._mesh_host)
            self.node_start = torch.from_numpy(st).to(self.dev)
            self.node_ref = torch.from_numpy(ref).to(self.dev)
            self.cell_partials = torch.zeros(4 * self.nc, dtype=torch.float32, device=self.dev)
            self.charge_f32 = torch.zeros(self.nn, dtype=torch.float32, device=self.dev)
            self.dep_bytes = int(lib().oppx_fempic_deposit_scratch_bytes(self.cap))
            self.dep_scratch = torch.zeros(self.dep_bytes, dtype=torch.uint8, device=self.dev)
        tail = (self.nc, self.node_start.data_ptr(), None, self.node_ref.data_ptr(), self.nn,
                self.cell_partials.data_ptr(), None, self.charge_f32.data_ptr(), self.dep_scratch.data_ptr(),
                self.dep_bytes, self.stream())
        if pending:
            check(lib().oppx_fempic_deposit_sorted_f32_perm(self.n_max, self.cap, self.counters.data_ptr(),
                                                            p2c.data_ptr(), lc.data_ptr(), self.perm.data_ptr(),
                                                            lc_f32.data_ptr(), *tail), "deposit_sorted_f32_perm")
        else:
            check(lib().oppx_fempic_deposit_sorted_f32(self.n_max, self.cap, self.counters.data_ptr(), p2c.data_ptr(),
                                                       lc.data_ptr(), *tail), "deposit_sorted_f32")
        return self.charge_f32

    def _rebind(self):
        self.parts = FempicParts(self.cap, self.pos.data_ptr(), self.vel.data_ptr(), self.lc.data_ptr(),
                                 self.p2c.data_ptr())
        self.dats = (DatDesc * 4)(DatDesc(self.pos.data_ptr(), 3, 8), DatDesc(self.vel.data_ptr(), 3, 8),
                                  DatDesc(self.lc.data_ptr(), 4, 8), DatDesc(self.p2c.data_ptr(), 1, 4))

    def deposit(self):
        self.materialize()
        check(lib().oppx_fempic_deposit_atomic(self.n_max, ctypes.byref(self.parts), ctypes.byref(self.mesh),
                                               self.charge.data_ptr(), self.counters.data_ptr(), self.stream()),
              "fempic_deposit")

    def step_begin(self):
        check(lib().oppx_fempic_step_begin(self.counters.data_ptr(), self.charge.data_ptr(), self.nn, self.stream()),
              "fempic_step_begin")

    def pass_fused(self, charge, mass, dt, record_hops=False, deposit=True, stage=False):
        """push + direct hop + walk (+ FP64 deposit). stage=True (for a step that a lazy sort follows,
        with the two-pass move): the new positions and velocities go to 48-byte records instead of the
        SoA dats, so that the step after the sort reads one record per particle through the permutation."""
        k = charge / mass * dt
        pending, staged = getattr(self, "pending", False), getattr(self, "staged", False)
        stage = stage and get_tuning()["inline_hops"] >= 1
        if staged and not pending:
            self.materialize()
            staged = False
        if stage or (pending and staged):
            self._pass_records(k, dt, record_hops, deposit, stage, pending)
            return
        if pending:  # the step reads its inputs through the pending sort
            swap = self._swap_buffers()
            # p2c None: the sort wrote the sorted p2c into the target already
            src = FempicParts(self.cap, self.pos.data_ptr(), self.vel.data_ptr(), self.lc.data_ptr(), None)
            dst = FempicParts(self.cap, swap[0].data_ptr(), swap[1].data_ptr(), swap[2].data_ptr(),
                              swap[3].data_ptr())
            check(lib().oppx_fempic_step_fused_gather(
                self.n_max, self.perm.data_ptr(), ctypes.byref(src), ctypes.byref(dst), ctypes.byref(self.mesh),
                self._dh_ref(), k, dt, self.counters.data_ptr(), self.hole_bits.data_ptr(), self.work.data_ptr(),
                self.work_bytes, self.hops.data_ptr() if record_hops else None,
                self.charge.data_ptr() if deposit else None, self.stream()), "fempic_step_fused_gather")
            cur = [self.pos, self.vel, self.lc, self.p2c]
            self.pos, self.vel, self.lc, self.p2c = swap
            self._swap = cur
            self._rebind()
            self.pending = False
            return
        check(lib().oppx_fempic_step_fused(self.n_max, ctypes.byref(self.parts), ctypes.byref(self.mesh),
                                           self._dh_ref(), k, dt, self.counters.data_ptr(),
                                           self.hole_bits.data_ptr(), self.work.data_ptr(), self.work_bytes,
                                           self.hops.data_ptr() if record_hops else None,
                                           self.charge.data_ptr() if deposit else None, self.stream()),
              "fempic_step_fused")

    def _pass_records(self, k, dt, record_hops, deposit, stage, pending):
        """The fused step through oppx_fempic_step_fused_rec: reads the staged records through the pending
        permutation (pending) and/or writes this step's records (stage)."""
        recs = self._records()
        rec_in = recs[self._rec_cur] if pending else None
        out_idx = (1 - self._rec_cur) if pending else self._rec_cur  # never the buffer being read
        rec_out = recs[out_idx] if stage else None
        if pending:
            swap = self._swap_buffers()
            dst = FempicParts(self.cap, swap[0].data_ptr(), swap[1].data_ptr(), swap[2].data_ptr(),
                              swap[3].data_ptr())
        else:
            dst = self.parts
        self._rec_call(self.perm.data_ptr() if pending else None, rec_in.data_ptr() if pending else None,
                       self.lc.data_ptr() if pending else None, dst,
                       rec_out.data_ptr() if rec_out is not None else None, k, dt,
                       self.hops.data_ptr() if record_hops else None, self.charge.data_ptr() if deposit else None)
        if pending:
            cur = [self.pos, self.vel, self.lc, self.p2c]
            self.pos, self.vel, self.lc, self.p2c = swap
            self._swap = cur
            self._rebind()
            self.pending = False
        if stage:
            self._rec_cur = out_idx
        self.staged = stage

    def _rec_call(self, perm, rec_in, src_lc, dst, rec_out, k, dt, hops, charge):
        check(lib().oppx_fempic_step_fused_rec(
            self.n_max, perm, rec_in, src_lc, ctypes.byref(dst), rec_out, ctypes.byref(self.mesh), self._dh_ref(),
            k, dt, self.counters.data_ptr(), self.hole_bits.data_ptr(), self.work.data_ptr(), self.work_bytes, hops,
            charge, self.stream()), "fempic_step_fused_rec")

    def step_fused(self, charge, mass, dt, record_hops=False, fill="holefill"):
        """push + direct hop + walk + deposit in one pass, then hole-fill (or a sort finalize)."""
        self.step_begin()
        self.pass_fused(charge, mass, dt, record_hops)
        if fill == "sort":
            self.sort(drop_removed=True)
        else:
            self.holefill()

    def step_unfused(self, charge, mass, dt, record_hops=True, fill="holefill"):
        """The reference's loop order: reset, push, move (+ hole-fill or sort finalize), deposit."""
        self.charge.zero_()
        self.push(charge, mass, dt)
        self.move(record_hops)
        if fill == "sort":
            self.sort(drop_removed=True)
        else:
            self.holefill()
        self.deposit()

    # -- readback ------------------------------------------------------------------------------
    def read_counters(self):
        c = self.counters.cpu().numpy()
        return dict(size=int(c[0]), removed=int(c[1]), holes_low=int(c[2]), max_hops=int(c[3]),
                    total_hops=int(c[4:6].view(np.int64)[0]), faults=int(c[6]))

    def sync_size(self):
        """Host copy of the live count (the D2H of opp_finalize_particle_move)."""
        self.n_max = self.read_counters()["size"]
        return self.n_max

    def download(self):
        self.materialize()
        n = self.read_counters()["size"]
        return dict(n=n, pos=self.pos[:, :n].T.cpu().numpy().copy(), vel=self.vel[:, :n].T.cpu().numpy().copy(),
                    lc=self.lc[:, :n].T.cpu().numpy().copy(), p2c=self.p2c[:n].cpu().numpy().copy(),
                    charge=self.charge.cpu().numpy().copy())


def concat_instances(instances) -> dict:
    """Host side of a batched ensemble: one mesh, field and particle set holding every instance,
    cell / node / p2c indices offset by the instance's base (c2c = -1 stays a boundary), the
    instance of every cell, and the overlays concatenated with offset cell indices."""
    if not instances:
        raise ValueError("empty batch")
    nc = np.array([int(d["mesh"]["n_cells"]) for d in instances], np.int64)
    nn = np.array([int(d["mesh"]["n_nodes"]) for d in instances], np.int64)
    if nc.sum() >= MAX_CELL_INDEX:
        raise ValueError("concatenated mesh too large for 32-bit cell indices")
    cb = np.concatenate([[0], np.cumsum(nc)[:-1]]).astype(np.int64)
    nb = np.concatenate([[0], np.cumsum(nn)[:-1]]).astype(np.int64)
    cat = np.concatenate
    for d, n in zip(instances, nc):
        if len(d["p2c"]) and not ((d["p2c"] >= 0) & (d["p2c"] < n)).all():
            raise ValueError("instance p2c out of its mesh")
    mesh = dict(
        n_cells=int(nc.sum()), n_nodes=int(nn.sum()),
        c_det=cat([d["mesh"]["c_det"].reshape(-1, 16) for d in instances]),
        c_vol=cat([d["mesh"]["c_vol"] for d in instances]),
        c2c=cat([np.where(d["mesh"]["c2c"] >= 0, d["mesh"]["c2c"] + b, d["mesh"]["c2c"]).astype(np.int32)
                 for d, b in zip(instances, cb)]),
        c2n=cat([(d["mesh"]["c2n"] + b).astype(np.int32) for d, b in zip(instances, nb)]),
    )
    ovs, offs, off = [], [], 0
    for d, b in zip(instances, cb):
        ov = d.get("overlay")
        if ov is None:
            offs.append(-1)
            continue
        if ov.size != int(np.prod(dh_dims(d["bmin"], d["bmax"], d["gs"]))):
            raise ValueError("overlay size does not match its grid dims")
        n_i = int(d["mesh"]["n_cells"])
        if not (((ov >= 0) & (ov < n_i)) | (ov == MAX_CELL_INDEX)).all():
            raise ValueError("overlay entries must be cells of the instance or MAX_CELL_INDEX")
        ovs.append(np.where(ov == MAX_CELL_INDEX, ov, ov + b).astype(np.int32))
        offs.append(off)
        off += ov.size
    return dict(mesh=mesh, ef=cat([d["ef"].reshape(-1, 3) for d in instances]),
                pos=cat([d["pos"] for d in instances]), vel=cat([d["vel"] for d in instances]),
                p2c=cat([(d["p2c"] + b).astype(np.int32) for d, b in zip(instances, cb)]),
                cell_base=cb, n_cells=nc, node_base=nb, n_nodes=nn,
                cell_inst=np.repeat(np.arange(len(instances), dtype=np.int32), nc),
                overlay=cat(ovs) if ovs else np.zeros(1, np.int32), overlay_off=np.array(offs, np.int64))


class FempicBatch(FempicDevice):
    """An ensemble of independent FEM-PIC instances resident on the GPU as ONE concatenated mesh and
    ONE particle set, stepped by single launches (oppx_fempic_batch_step_fused, the sort, the FP32
    deposit) that cover every instance.

    instances: list of dicts with mesh, ef, pos, vel, p2c, charge, mass, dt and optionally overlay,
    bmin, bmax, gs (direct hop). Cell, node and p2c indices are offset by each instance's base; the
    Sort_All finalize (stable by concatenated cell) keeps each instance's particles contiguous and in
    the order the instance would have alone, so every per-instance result equals its alone run.
    """

    def __init__(self, instances, device="cuda"):
        import torch
        u = concat_instances(instances)
        super().__init__(u["mesh"], u["ef"], u["pos"], u["vel"], u["p2c"], device=device)
        self.n_inst = len(instances)
        self.cell_base, self.n_cells_i = u["cell_base"], u["n_cells"]
        self.node_base, self.n_nodes_i = u["node_base"], u["n_nodes"]
        self.cell_inst = torch.from_numpy(u["cell_inst"]).to(self.dev)
        self.overlay_cat = torch.from_numpy(u["overlay"]).to(self.dev)
        table = (BatchInst * self.n_inst)()
        for i, d in enumerate(instances):
            t = table[i]
            t.cell_base, t.n_cells = int(u["cell_base"][i]), int(u["n_cells"][i])
            t.node_base, t.n_nodes = int(u["node_base"][i]), int(u["n_nodes"][i])
            if u["overlay_off"][i] >= 0:
                dims = dh_dims(d["bmin"], d["bmax"], d["gs"])
                t.dh_cell = self.overlay_cat.data_ptr() + 4 * int(u["overlay_off"][i])
                for a in range(3):
                    t.dims[a] = int(dims[a])
                    t.min[a] = float(d["bmin"][a])
                t.inv_gs = 1.0 / float(d["gs"])
            t.dt = float(d["dt"])
            t.k = float(d["charge"]) / float(d["mass"]) * float(d["dt"])
        raw = np.frombuffer(bytes(table), np.uint8).copy()
        self.inst_table = torch.from_numpy(raw).to(self.dev)

    def pass_batch(self, deposit=False, record_hops=False, stage=False):
        """push + direct hop + walk (+ FP64 deposit into the concatenated node charge) of every instance.
        stage=True (with the two-pass move): the new positions and velocities go to 48-byte records, and a
        pass after a lazy sort of staged records reads them through the permutation, as
        FempicDevice.pass_fused does for one instance."""
        pending, staged = self.pending, self.staged
        stage = stage and get_tuning()["inline_hops"] >= 1
        if staged and not pending:
            self.materialize()
            staged = False
        if stage or (pending and staged):
            self._pass_records(0.0, 0.0, record_hops, deposit, stage, pending)
            return
        self.materialize()
        check(lib().oppx_fempic_batch_step_fused(
            self.n_max, ctypes.byref(self.parts), ctypes.byref(self.mesh), self.cell_inst.data_ptr(),
            self.inst_table.data_ptr(), self.n_inst, self.counters.data_ptr(), self.hole_bits.data_ptr(),
            self.work.data_ptr(), self.work_bytes, self.hops.data_ptr() if record_hops else None,
            self.charge.data_ptr() if deposit else None, self.stream()), "fempic_batch_step_fused")

    def _rec_call(self, perm, rec_in, src_lc, dst, rec_out, k, dt, hops, charge):
        # k and dt come from the instance table
        check(lib().oppx_fempic_batch_step_fused_rec(
            self.n_max, perm, rec_in, src_lc, ctypes.byref(dst), rec_out, ctypes.byref(self.mesh),
            self.cell_inst.data_ptr(), self.inst_table.data_ptr(), self.n_inst, self.counters.data_ptr(),
            self.hole_bits.data_ptr(), self.work.data_ptr(), self.work_bytes, hops, charge, self.stream()),
            "fempic_batch_step_fused_rec")

    def step_batch(self, deposit="f32"):
        """One ensemble step: reset, batched push/move, Sort_All finalize, deposit.
        deposit: "f32" (compensated FP32 over the sorted particles, into self.charge_f32) or "f64"
        (FP64 atomics inside the walk, into self.charge)."""
        self.step_begin()
        self.pass_batch(deposit=(deposit == "f64"))
        self.sort(drop_removed=True)
        if deposit == "f32":
            self.deposit_sorted_f32()

    def steps_batch(self, steps, deposit="f32", lazy=True, mark=None):
        """`steps` ensemble steps, each one step_batch's work, ending with the particle dats sorted in place.
        lazy (with the two-pass move): every sort computes only the permutation and the sorted p2c; each
        pass writes positions and velocities as 48-byte records and the next pass reads them through the
        permutation; the FP32 deposit reads the sorted p2c and a gathered lc; after the last step the
        permutation is applied (materialize). The particle order, cells, positions, velocities, lc and
        charges after each step equal step_batch's. An even `steps` ends on the starting buffers, so the
        sequence can be captured in a graph and replayed. mark(phase) (optional) is called before each
        step (0), after its begin and pass (1), after its sort (2) and after its deposit (3)."""
        lazy = lazy and get_tuning()["inline_hops"] >= 1
        note = mark or (lambda phase: None)
        for _ in range(steps):
            note(0)
            self.step_begin()
            self.pass_batch(deposit=(deposit == "f64"), stage=lazy)
            note(1)
            self.sort(drop_removed=True, lazy=lazy)
            note(2)
            if deposit == "f32":
                self.deposit_sorted_f32()
            note(3)
        self.materialize()

    def split(self):
        """Per-instance state (host): the concatenated particles are sorted by concatenated cell."""
        out = self.download()
        f32 = self.charge_f32.cpu().numpy() if hasattr(self, "charge_f32") else None
        res = []
        for i in range(self.n_inst):
            c0, c1 = int(self.cell_base[i]), int(self.cell_base[i] + self.n_cells_i[i])
            lo, hi = np.searchsorted(out["p2c"], [c0, c1])
            n0, n1 = int(self.node_base[i]), int(self.node_base[i] + self.n_nodes_i[i])
            res.append(dict(n=int(hi - lo), pos=out["pos"][lo:hi], vel=out["vel"][lo:hi], lc=out["lc"][lo:hi],
                            p2c=out["p2c"][lo:hi] - c0, charge=out["charge"][n0:n1],
                            charge_f32=None if f32 is None else f32[n0:n1]))
        return res


class NesoDevice:
    """NESO-advection particles + periodic quad mesh resident on the GPU (C-ABI driven)."""

    def __init__(self, mesh: dict, pos: np.ndarray, vel: np.ndarray, p2c: np.ndarray, mdir=None, device="cuda",
                 grid: bool = True):
        """grid=True: when the mesh is, bit for bit, the nx x ny periodic grid, the walk computes the
        visited records instead of gathering them (same cells, hops and positions)."""
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        L = lib()
        check(L.oppx_init(self.dev.index or 0), "oppx_init")
        nc = mesh["n_cells"]
        if not ((mesh["c2c"] >= 0).all() and (mesh["c2c"] < nc).all()):
            raise ValueError("c2c entries out of [0, n_cells)")
        if pos.shape[0] and not ((p2c >= 0).all() and (p2c < nc).all()):
            raise ValueError("p2c entries out of [0, n_cells)")
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(self.dev, dtype=dt)
        self.nc = nc
        self.nx, self.ny, self.cw = mesh["nx"], mesh["ny"], mesh["cell_width"]
        self.ll = t(mesh["c_pos_ll"].reshape(-1), torch.float64)
        self.c2c = t(mesh["c2c"].reshape(-1), torch.int32)
        self.cells = torch.zeros(nc * 4, dtype=torch.float64, device=self.dev)  # 32 B per quad
        check(L.oppx_quad_cells_build(self.ll.data_ptr(), self.c2c.data_ptr(), nc, self.cells.data_ptr(),
                                      self.stream()), "quad_cells_build")
        self.grid = False
        if grid and self.nx * self.ny == nc:
            bad = torch.zeros(1, dtype=torch.int32, device=self.dev)
            check(L.oppx_quad_grid_check(self.cells.data_ptr(), nc, self.nx, self.ny, self.cw, bad.data_ptr(),
                                         self.stream()), "quad_grid_check")
            self.grid = int(bad.item()) == 0
        n = pos.shape[0]
        self.n = n
        cap = max(n, 1)
        self.pos = torch.zeros((2, cap), dtype=torch.float64, device=self.dev)
        self.vel = torch.zeros((2, cap), dtype=torch.float64, device=self.dev)
        self.mdir = torch.ones((2, cap), dtype=torch.int32, device=self.dev)
        self.p2c = torch.zeros(cap, dtype=torch.int32, device=self.dev)
        if n:
            self.pos[:, :n] = t(pos.T, torch.float64)
            self.vel[:, :n] = t(vel.T, torch.float64)
            self.p2c[:n] = t(p2c, torch.int32)
            if mdir is not None:
                self.mdir[:, :n] = t(mdir.T, torch.int32)
        self.parts = NesoParts(cap, self.pos.data_ptr(), self.vel.data_ptr(), self.mdir.data_ptr(), self.p2c.data_ptr())
        self.counters = torch.zeros(COUNTER_WORDS, dtype=torch.int32, device=self.dev)
        self.counters[0] = n
        self.hops = torch.zeros(cap, dtype=torch.int32, device=self.dev)

    def stream(self):
        return self.torch.cuda.current_stream(self.dev).cuda_stream

    def step(self, dt, record_hops=False, do_update=True):
        flags = (1 if do_update else 0) | 2
        hops = self.hops.data_ptr() if record_hops else None
        if self.grid:
            check(lib().oppx_neso_step_grid(self.n, ctypes.byref(self.parts), self.nx, self.ny, dt, self.cw,
                                            self.nx * self.cw, self.ny * self.cw, flags, self.counters.data_ptr(), hops,
                                            self.stream()), "neso_step_grid")
            return
        check(lib().oppx_neso_step(self.n, ctypes.byref(self.parts), self.cells.data_ptr(), self.nc, dt, self.cw,
                                   self.nx * self.cw, self.ny * self.cw, flags, self.counters.data_ptr(), hops,
                                   self.stream()), "neso_step")

    def sort(self):
        """Stable reorder of the particle dats by cell (opp_particle_sort): the moved particles' cell
        records are then read by neighbouring lanes. The swap buffers come in on the first call."""
        torch = self.torch
        if not hasattr(self, "_swap"):
            cap = self.p2c.shape[0]
            self._swap = [torch.empty_like(self.pos), torch.empty_like(self.vel), torch.empty_like(self.mdir),
                          torch.empty_like(self.p2c)]
            self.sort_scratch_bytes = int(lib().oppx_sort_scratch_bytes(cap))
            self.sort_scratch = torch.zeros(self.sort_scratch_bytes, dtype=torch.uint8, device=self.dev)
        cur = [self.pos, self.vel, self.mdir, self.p2c]
        dims = [(2, 8), (2, 8), (2, 4), (1, 4)]
        desc = lambda ts: (DatDesc * 4)(*[DatDesc(t.data_ptr(), d, e) for t, (d, e) in zip(ts, dims)])
        check(lib().oppx_sort_by_cell(self.n, self.p2c.shape[0], self.counters.data_ptr(), self.nc, self.p2c.data_ptr(),
                                      desc(cur), desc(self._swap), 4, None, 0, self.sort_scratch.data_ptr(),
                                      self.sort_scratch_bytes, self.stream()), "sort_by_cell")
        self.pos, self.vel, self.mdir, self.p2c = self._swap
        self._swap = cur
        self.parts = NesoParts(self.p2c.shape[0], self.pos.data_ptr(), self.vel.data_ptr(), self.mdir.data_ptr(),
                               self.p2c.data_ptr())

    def read_counters(self):
        c = self.counters.cpu().numpy()
        return dict(size=int(c[0]), max_hops=int(c[3]), total_hops=int(c[4:6].view(np.int64)[0]), faults=int(c[6]))

    def download(self):
        n = self.n
        return dict(pos=self.pos[:, :n].T.cpu().numpy().copy(), mdir=self.mdir[:, :n].T.cpu().numpy().copy(),
                    p2c=self.p2c[:n].cpu().numpy().copy(), hops=self.hops[:n].cpu().numpy().copy())
