"""Writes the fixture of OP-PIC's own generated NESO-advection main run unchanged, from the REFERENCE
build (oracle/_ref/advec_ref_seq: apps_codegen/app_neso_advection_mdir_cg/advec_opp.cpp + the generated
seq loops + the seq backend, opp_exit wrapped by tests/api/app_dump_at_exit.cpp).

Run where the reference tree exists:  python tests/golden/make_neso_app_golden.py
The main builds its own mesh and particles (std::mt19937 seeds 52234234 / 52234231) from the deck
tests/golden/neso_app_small.param. Stored: the per-step verify error count and particle count the main
logs, and the final particle dats. Fixture is .npz (allow_pickle=False).
"""

from __future__ import annotations

import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
GOLDEN = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(ROOT, "oracle", "_ref", "advec_ref_seq")
PARAMS = os.path.join(GOLDEN, "neso_app_small.param")

STEP_RE = re.compile(r"^Main\[0\]\[\d+\] - ts: (\d+) \| errors: (-?\d+) \| .*Gbl parts: (\d+) ")


def parse_log(text):
    """(step, errors, parts) of every step line the generated main prints."""
    rows = [tuple(int(g) for g in m.groups()) for m in map(STEP_RE.match, text.splitlines()) if m]
    return np.array(rows, dtype=np.int64).reshape(-1, 3)


def read_dump(d, n):
    """Final particle dats written by the opp_exit hook (host AoS order)."""
    f = lambda name, dt: np.fromfile(os.path.join(d, name + ".bin"), dtype=dt)
    pos, vel = f("p_pos", np.float64).reshape(-1, 2), f("p_vel", np.float64).reshape(-1, 2)
    mdir, p2c, idx = f("p_move_dir", np.int32).reshape(-1, 2), f("p_mesh_rel", np.int32), f("p_index", np.int32)
    assert len(pos) == len(vel) == len(mdir) == len(p2c) == len(idx) == n
    return dict(pos=pos, vel=vel, mdir=mdir, p2c=p2c, index=idx)


def main():
    with tempfile.TemporaryDirectory() as td:
        r = subprocess.run([REF, PARAMS], capture_output=True, text=True, cwd=td, env=dict(os.environ, OPPX_APP_DUMP=td))
        if r.returncode != 0:
            raise SystemExit(f"reference app failed rc={r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        steps = parse_log(r.stdout)
        assert len(steps) == 30 and (steps[:, 0] == np.arange(30)).all()
        out = read_dump(td, int(steps[-1, 2]))
    np.savez_compressed(os.path.join(GOLDEN, "neso_app_small.npz"), steps=steps, **out)
    print("neso_app_small:", len(out["p2c"]), "particles, errors per step", steps[:, 1].tolist())


if __name__ == "__main__":
    main()
