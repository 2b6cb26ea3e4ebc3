"""The reference's batch example end to end through the command-line flow (vdock.driver, the
restatement of V11/main/main.cpp's main / main_procedure): the ligand folders of
QVW/input_file_example/auto_test/{1u1b,2bm2,2er7}/ (ten or eleven ligands each, fixtures of
test_oracle_rescore_qvw_auto.py) docked in --ligand_directory mode with the boxes of
auto_test/<receptor>_config.txt and thread = 16384, as the reference ran them, for seeds 0 and 1.

Checked:
  - one output per ligand under the reference's name (<output_directory>/<ligand>_out.pdbqt),
    with up to nine modes, sorted, the first at rmsd 0;
  - the written files rescored by the oracle give back their own printed scores (0.05 of the
    printing + 0.01 for the coordinates' rounding) and rmsd bounds (2e-3): the writer, the final
    scores and the mode selection agree with the FP64 restatement on fresh outputs.  The scores'
    offset is the intra energy of the first pose by refined energy (V11/main/main.cpp:439-440);
    the refined energy (smooth tables, FP64 refinement) differs from the total rescored from the
    printed coordinates by up to 0.014 here and the rounding of the coordinates adds about 0.01,
    so in a near tie that pose is any of those within 0.05 of the lowest rescored total;
  - the search finds minima as deep as the reference's: over the 31 ligands, the best score of
    the two seeds is on average no higher than the reference's written best, and for each ligand
    at most 0.75 higher.  Docking is stochastic: the reference's own three results for the 1u1b
    ligand on its receptor (V11 example -7.9, QuickVina-W example -7.0, auto_test -6.2) span
    1.7 kcal/mol;
  - docking the first folder twice with the same seed, each time in a fresh context, writes the
    same files byte for byte.

Every run docks in a context of its own, as every run of the reference starts a process of its
own: ligand ids, which key the random streams, count from 0 in each, so the outcome does not
depend on which tests ran before."""
import os
import shutil

import numpy as np
import pytest

import oracle as orc
from conftest import GOLDEN
from test_oracle_rescore import parse_models
from test_oracle_rescore_qvw_auto import BOXES, CASES, FX, grid

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def check_written(rec, text, lig):
    """Printed scores and rmsd bounds of a written output against the oracle's rescoring."""
    models, remarks = parse_models(text)
    assert 1 <= len(models) <= 9
    ser = lig.serials()
    X = [np.array([m[s] for s in ser]) for m in models]
    sc = [r[0] for r in remarks]
    assert sc == sorted(sc) and remarks[0][1:] == (0.0, 0.0)
    g = grid(rec)
    inter = np.array([orc.noncache_eval(g, lig, x) for x in X])
    intra = np.array([orc.intra_fast(g, lig, x) for x in X])
    tot = inter + intra
    # the offset is the intra energy of the first pose by refined energy; rescored from printed
    # coordinates, that pose is one of those within 0.05 of the lowest total
    dev = [np.abs([lig.conf_independent(e - intra[j]) for e in tot] - np.array(sc)).max()
           for j in np.flatnonzero(tot <= tot.min() + 0.05)]
    assert min(dev) <= 0.06, (dev, sc)
    for x, (_, lb_ref, ub_ref) in zip(X, remarks):
        lb, ub = lig.rmsd_bounds(x, X[0])
        assert abs(lb - lb_ref) <= 2e-3 and abs(ub - ub_ref) <= 2e-3
    return sc[0]


def _run_fresh(driver, opts):
    import vdock
    fresh = vdock.Context(0)
    try:
        return driver.run(opts, ctx=fresh, log=lambda *a: None)
    finally:
        fresh.close()


def test_ligand_directory_flow_on_auto_test(tmp_path):
    from vdock import driver
    diffs = []
    seed0_1u1b = None
    for rec in ("1u1b", "2bm2", "2er7"):
        d = tmp_path / rec
        d.mkdir()
        cases = {name: k for r, name, k in CASES if r == rec}
        for name in cases:
            src = os.path.join(GOLDEN, "1u1b_ligand.pdbqt") if name == "1u1b_ligand" else \
                os.path.join(GOLDEN, "qvw_auto", name + ".pdbqt")
            shutil.copyfile(src, d / (name + ".pdbqt"))
        (c, s) = BOXES[rec]
        best = {}
        for seed in (0, 1):
            out_dir = str(tmp_path / f"{rec}_out{seed}")
            opts = driver.resolve(receptor=os.path.join(GOLDEN, rec + "_protein.pdbqt"), ligand_directory=str(d),
                                  output_directory=out_dir, thread=16384, seed=seed, center_x=c[0], center_y=c[1],
                                  center_z=c[2], size_x=s[0], size_y=s[1], size_z=s[2])
            say(rec, "seed", seed, "docking", len(cases), "ligands")
            res = _run_fresh(driver, opts)
            assert len(res) == len(cases)
            if rec == "1u1b" and seed == 0:
                seed0_1u1b = {os.path.basename(r["ligand"]): r["text"] for r in res}
            for r in res:
                name = os.path.basename(r["ligand"])[:-len(".pdbqt")]
                assert r["out"] == os.path.join(out_dir, name + "_out.pdbqt")
                with open(r["out"]) as f:
                    text = f.read()
                assert text == r["text"]
                top = check_written(rec, text, orc.Ligand((d / (name + ".pdbqt")).read_text()))
                best[name] = min(best.get(name, 1e9), top)
        for name, k in cases.items():
            ref = FX[f"remark_{k}"][0][0]
            say(rec, name, "best of two seeds", best[name], "reference", ref)
            diffs.append(best[name] - ref)
    # the same inputs and seed in another fresh context give the same files
    c, s = BOXES["1u1b"]
    res = _run_fresh(driver, driver.resolve(receptor=os.path.join(GOLDEN, "1u1b_protein.pdbqt"),
                                            ligand_directory=str(tmp_path / "1u1b"), output_directory=str(tmp_path / "again"),
                                            thread=16384, seed=0, center_x=c[0], center_y=c[1], center_z=c[2],
                                            size_x=s[0], size_y=s[1], size_z=s[2]))
    again = {os.path.basename(r["ligand"]): r["text"] for r in res}
    assert len(again) == 11 and again == seed0_1u1b
    diffs = np.array(diffs)
    say("best minus reference: mean", diffs.mean(), "max", diffs.max())
    assert len(diffs) == 31
    assert diffs.mean() <= 0.0 and diffs.max() <= 0.75
