"""The C-ABI library loads and exports every symbol include/oppx.h declares (CPU only: no
compute call needs a GPU here), and the host builders behave."""

import ctypes
import os
import re

import numpy as np

import oppx


def header_functions(root):
    text = open(os.path.join(root, "include", "oppx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s+(oppx_[a-z0-9_]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(root):
    names = header_functions(root)
    assert len(names) >= 20
    L = ctypes.CDLL(oppx.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding declares a prototype for each of them
    assert sorted(set(names) - set(oppx.declared_symbols())) == []


def test_error_path_reports_message():
    rc = oppx.lib().oppx_box_mesh_sizes(0, 1, 1, ctypes.byref(ctypes.c_int64()), ctypes.byref(ctypes.c_int64()))
    assert rc == oppx.ERR_ARG


def test_overlay_out_of_cuboid_and_dims():
    # a 32^3 overlay on [0,1]^3 has 33 boxes per side (floating step count + 1)
    assert list(oppx.dh_dims((0, 0, 0), (1, 1, 1), 1 / 32)) == [33, 33, 33]
    assert list(oppx.dh_dims((0, 0, 0), (1, 1, 1), 1 / 128)) == [129, 129, 129]
    m = oppx.box_mesh(4, 4, 4)
    ov = oppx.dh_overlay(m, (0, 0, 0), (1, 1, 1), 1 / 8)
    assert ((ov >= 0) & (ov < m["n_cells"])).all()


def test_neso_mesh_periodic():
    m = oppx.neso_mesh(5, 4, 0.5)
    c2c = m["c2c"]
    for i in range(20):
        x, y = i % 5, i // 5
        assert c2c[i, 0] == (x - 1) % 5 + 5 * y and c2c[i, 1] == (x + 1) % 5 + 5 * y
        assert c2c[i, 2] == x + 5 * ((y - 1) % 4) and c2c[i, 3] == x + 5 * ((y + 1) % 4)
    assert np.array_equal(m["c_pos_ll"][7], [2 * 0.5, 1 * 0.5])


OPP_API = [
    "opp_init(int, char**)", "opp_exit()", "opp_decl_set(int, char const*)",
    "opp_decl_particle_set(char const*, opp_set_core*)", "opp_decl_particle_set(int, char const*, opp_set_core*)",
    "opp_decl_map(opp_set_core*, opp_set_core*, int, int*, char const*)",
    "opp_decl_dat(opp_set_core*, int, opp_data_type, void*, char const*)",
    "opp_decl_const_impl(int, int, char*, char const*)", "opp_increase_particle_count(opp_set_core*, int)",
    "opp_reset_dat_impl(opp_dat_core*, char*, opp_reset)", "opp_particle_sort(opp_set_core*)",
    "opp_init_particle_move(opp_set_core*, int, opp_arg*)", "opp_finalize_particle_move(opp_set_core*)",
    "opp_par_loop_all__calculate_new_pos_vel_kernel(opp_set_core*, opp_arg, opp_arg, opp_arg)",
    "opp_particle_move__move_kernel(opp_set_core*, opp_map_core*, opp_map_core*, opp_arg, opp_arg, opp_arg, opp_arg)",
    "opp_par_loop_all__deposit_charge_on_nodes_kernel(opp_set_core*, opp_arg, opp_arg, opp_arg, opp_arg, opp_arg)",
    "opp_init_direct_hop_cg(double, opp_dat_core*, opp::BoundingBox const&, opp_map_core*, opp_map_core*, opp_arg, "
    "opp_arg, opp_arg, opp_arg)",
    "opp_par_loop_all__update_pos_kernel(opp_set_core*, opp_arg, opp_arg)",
    "opp_particle_move__move_kernel(opp_set_core*, opp_map_core*, opp_map_core*, opp_arg, opp_arg, opp_arg)",
]


def test_opp_lib_dropin_exports_api(root):
    """libopp_mi355x.so defines the opp_lib entry points OP-PIC's generated FEM-PIC and NESO mains call."""
    import subprocess

    so = os.path.join(root, "op-pic_amd", "libopp_mi355x.so")
    out = subprocess.run(["nm", "-DC", "--defined-only", so], check=True, capture_output=True, text=True).stdout
    defined = {line.split(" ", 2)[2] for line in out.splitlines() if " T " in line}
    missing = [s for s in OPP_API if s not in defined]
    assert not missing, missing
    # the drop-in links against the C-ABI library, not against any torch or python runtime
    deps = subprocess.run(["ldd", so], capture_output=True, text=True).stdout
    assert "liboppx.so" in deps and "torch" not in deps and "python" not in deps
