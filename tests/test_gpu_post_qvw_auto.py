"""The step after the search on the device for the QuickVina-W example's 31 batch outputs
(QVW/input_file_example/auto_test/{1u1b,2bm2,2er7}/, fixtures of test_oracle_rescore_qvw_auto.py):
per receptor, one batched post-processing launch over all of its ligands (ten or eleven, 13 to
89 atoms), each with the nine written modes as its poses -- FP64 refinement against the receptor
atoms, final scores, mode selection and rmsd bounds -- against the oracle's post-processing of
the same poses.

Scores and rmsd bounds agree to 1e-3 per mode.  A refinement is a BFGS run whose line search
branches on comparisons of energies, so the last bits of the device's and the host's sine,
cosine and sums can send one pose down another path; a ligand whose modes differ that way is
accepted only for at most two ligands of 31, with its best score still within 0.05, and is
reported."""
import os

import numpy as np
import pytest

import oracle as orc
from conftest import GOLDEN, golden
from test_oracle_rescore_qvw_auto import BOXES, CASES, ligand_text, modes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_batched_postprocess_of_auto_test_outputs(ctx):
    import vdock
    branched = []
    for rec in ("1u1b", "2bm2", "2er7"):
        ctx.set_receptor_pdbqt(golden(rec + "_protein.pdbqt"))
        box = vdock.box_from_center(*BOXES[rec], 0.375)
        say(rec, "populate", box.dims)
        ctx.grid_populate(box)
        xyz, xs = orc.parse_receptor(golden(rec + "_protein.pdbqt"))[:2]
        g = orc.Grid(orc.Lut(), *box.as_tuple())
        g.set_receptor(xyz, xs)
        cases = [(name, k) for r, name, k in CASES if r == rec]
        fitted = np.load(os.path.join(GOLDEN, "qvw_auto_confs.npz"), allow_pickle=False)
        ligs, ids, confs = [], [], []
        for name, k in cases:
            text = ligand_text(name)
            lig = orc.Ligand(text)
            ligs.append(lig)
            ids.append(ctx.add_ligand(text))
            # FP32 conformations fitted to the written modes (both sides start from these values)
            c = fitted[f"conf_{k}"].astype(np.float64)
            assert max(np.abs(lig.set_conf(ci) - x).max() for ci, x in zip(c, modes(k, lig))) < 0.02
            confs.append(c)
        say(rec, "postprocess", len(ids), "ligands")
        st = ctx.postprocess(ids, confs)
        assert st.blocks == 9 * len(ids)
        for j, ((name, _), lig) in enumerate(zip(cases, ligs)):
            dm, _, _ = ctx.post_results(j)
            om = orc.postprocess(g, lig, confs[j], orc.refine_steps(lig))
            de = [abs(m.e - o["e"]) for m, o in zip(dm, om)]
            dr = [max(abs(m.rmsd_lb - o["lb"]), abs(m.rmsd_ub - o["ub"])) for m, o in zip(dm, om)]
            same = len(dm) == len(om) and max(de) <= 1e-3 and max(dr) <= 1e-3
            say(rec, name, "modes", len(dm), len(om), "max |de|", max(de), "max |drmsd|", max(dr))
            if not same:
                branched.append((rec, name))
                assert abs(dm[0].e - om[0]["e"]) <= 0.05, (rec, name, dm[0].e, om[0]["e"])
    say("ligands whose refinement took another path:", branched)
    assert len(branched) <= 2
