"""The compiled C++ consumer of include/vdock.h (vina-cuda_amd/examples/vdock_dock: the
reference's --ligand_directory run, main_procedure_cl with the output containers that cl_to_vina
fills, and the post-search block of main_procedure) on the 1u1b folder of the QuickVina-W
auto_test example (eleven ligands, the example's box, thread = 16384, seed 0):

  - it writes one <ligand>_out.pdbqt per ligand, equal byte for byte to what the Python driver
    (vdock.driver, the same flow through ctypes) writes for the same options in a fresh context;
  - each ligand's pose list (the output container: kept chains and the best energy) equals the
    one vd_batch_results + vd_cluster give for the same batch seen from Python."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from helpers import say
from test_oracle_rescore_qvw_auto import BOXES, CASES

pytestmark = pytest.mark.gpu


def test_cpp_consumer_matches_python_driver(tmp_path):
    import vdock
    from vdock import driver
    exe = os.path.join(ROOT, "vina-cuda_amd", "examples", "vdock_dock")
    d = tmp_path / "1u1b"
    d.mkdir()
    names = sorted(name for r, name, _ in CASES if r == "1u1b")
    for name in names:
        src = os.path.join(GOLDEN, "1u1b_ligand.pdbqt") if name == "1u1b_ligand" else \
            os.path.join(GOLDEN, "qvw_auto", name + ".pdbqt")
        shutil.copyfile(src, d / (name + ".pdbqt"))
    (c, s) = BOXES["1u1b"]
    rec = os.path.join(GOLDEN, "1u1b_protein.pdbqt")
    opts = dict(center_x=c[0], center_y=c[1], center_z=c[2], size_x=s[0], size_y=s[1], size_z=s[2], thread=16384, seed=0)
    args = [exe, "--receptor", rec, "--ligand_directory", str(d), "--output_directory", str(tmp_path / "cpp")]
    for k, v in opts.items():
        args += ["--" + k, str(v)]
    say("C++ consumer on", len(names), "ligands")
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    say(r.stdout[-400:], r.stderr[-400:])
    assert r.returncode == 0, r.stderr
    fresh = vdock.Context(0)
    try:
        res = driver.run(driver.resolve(receptor=rec, ligand_directory=str(d), output_directory=str(tmp_path / "py"), **opts),
                         ctx=fresh, log=lambda *a: None)
    finally:
        fresh.close()
    assert len(res) == len(names)
    for name in names:
        with open(tmp_path / "cpp" / (name + "_out.pdbqt")) as f:
            cpp = f.read()
        with open(tmp_path / "py" / (name + "_out.pdbqt")) as f:
            py = f.read()
        assert cpp == py and cpp.startswith("MODEL 1"), name
    # the output containers: vd_batch_results + vd_cluster of the same batch, from Python
    ctx = vdock.Context(0)
    try:
        ctx.set_receptor_pdbqt(open(rec).read())
        ctx.grid_populate(vdock.box_from_center(list(c), list(s), 0.375))
        ids = [ctx.add_ligand((d / (n + ".pdbqt")).read_text()) for n in names]
        ctx.dock_batch(ids, vdock.default_params(chains=16384, seed=0))
        lines = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("pose list ")]
        assert len(lines) == len(names)
        for k, ln in enumerate(lines):
            e, _, heavy = ctx.batch_results(k)
            kept = vdock.cluster(e, heavy, 1.0, 20)
            assert ln[2] == names[k] + ".pdbqt" and int(ln[3]) == len(kept)
            assert abs(float(ln[4]) - float(e[kept[0]])) <= 5e-7 * max(1.0, abs(float(e[kept[0]])))
    finally:
        ctx.close()
