"""OP-PIC's generated application code compiles unchanged against the drop-in's headers
(op-pic_amd/include) and links with libopp_mi355x.so. CPU only; skipped where the OP-PIC source tree
is absent. The GPU run of the linked NESO main is tests/test_gpu_api.py."""

import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
INC = os.path.join(ROOT, "op-pic_amd", "include")
PKG = os.path.join(ROOT, "op-pic_amd")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "apps_codegen")),
                                reason="needs the OP-PIC source tree")


def gxx(*args):
    r = subprocess.run(["g++", "-std=c++17", "-w", "-I" + INC, *args], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def test_neso_generated_main_compiles_and_links_unchanged():
    app = os.path.join(REF, "apps_codegen", "app_neso_advection_mdir_cg")
    with tempfile.TemporaryDirectory() as td:
        gxx("-O1", "-I" + app, os.path.join(app, "advec_opp.cpp"), "-o", os.path.join(td, "advec"), "-L" + PKG,
            "-lopp_mi355x", "-loppx", "-Wl,-rpath," + PKG)


@pytest.mark.parametrize("header", ["fempic_misc_mesh_loader.h", "fempic_misc.h", "fempic_defs.h"])
def test_fempic_app_headers_compile(header):
    # fempic_opp.cpp itself includes the PETSc field solver (field_solver.h), out of scope here
    app = os.path.join(REF, "apps_codegen", "app_fempic_cg")
    gxx("-fsyntax-only", "-I" + app, "-x", "c++", os.path.join(app, header))


def test_cabana_generated_main_compiles_unchanged():
    # compiles; linking needs CabanaPIC's field loops, which the library does not serve
    app = os.path.join(REF, "apps_codegen", "app_cabanapic_cg")
    gxx("-fsyntax-only", "-I" + app, os.path.join(app, "cabana_opp.cpp"))
