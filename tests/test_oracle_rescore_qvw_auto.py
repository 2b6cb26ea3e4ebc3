"""The QuickVina-W example's batch results as anchors for the oracle's final scoring (CPU only):
the 31 outputs under QVW/input_file_example/auto_test/{1u1b,2bm2,2er7}/ -- ten DrugBank ligands
and the 1u1b ligand on 1u1b, and ten ligands each on 2bm2 and 2er7, nine modes each, docked by
the reference in --ligand_directory mode with the boxes of auto_test/{1u1b,2bm2,2er7}_config.txt
(fixtures tests/golden/qvw_auto_test.npz and tests/golden/qvw_auto/, written by
tests/golden/make_qvw_auto_test.py).

Each mode's REMARK VINA RESULT holds the score eval_adjusted gave it and its rmsd bounds to the
best mode (QVW/main/main.cpp:431-450, as V11/main/main.cpp:440-470).  The intra-molecular offset
of eval_adjusted is the intra energy of the first pose after sorting by the refined energy
(main.cpp:433-434), taken before the second sort, so where two modes tie on energy it can be the
intra energy of the mode printed second: DB04226 on 2bm2 is such a case (its two best modes differ
by 7e-4 kcal/mol in total energy, by 0.3 in intra energy), and it reproduces only with that rule.
Rescoring the printed coordinates must give the printed numbers back: rmsd bounds within 2e-3
(coordinates printed to 1e-3 A); scores within the 0.05 of their printing to one decimal plus 0.01
for the coordinates' rounding (moving every coordinate of these modes by up to +-5e-4 A moves
their scores by at most 0.007)."""
import os

import numpy as np
import pytest

import oracle as orc
from conftest import GOLDEN, golden

BOXES = {   # QVW/input_file_example/auto_test/<receptor>_config.txt
    "1u1b": ([-30.661, -29.540, 17.010], [24.0, 26.25, 22.5]),
    "2bm2": ([40.415, 110.986, 82.673], [29.0, 29.0, 29.0]),
    "2er7": ([2.865, 30.924, 17.484], [50.0, 26.25, 22.5]),
}

FX = np.load(os.path.join(GOLDEN, "qvw_auto_test.npz"), allow_pickle=False)
CASES = [(str(r), str(l), k) for k, (r, l) in enumerate(zip(FX["receptor"], FX["ligand"]))]


def ligand_text(name):
    if name == "1u1b_ligand":
        return golden("1u1b_ligand.pdbqt")
    with open(os.path.join(GOLDEN, "qvw_auto", name + ".pdbqt")) as f:
        return f.read()


_grids = {}


def grid(rec):
    if rec not in _grids:
        xyz, xs = orc.parse_receptor(golden(rec + "_protein.pdbqt"))[:2]
        g = orc.Grid(orc.Lut(), *orc.grid_dims(*BOXES[rec]))
        g.set_receptor(xyz, xs)
        _grids[rec] = g
    return _grids[rec]


def modes(k, lig):
    """Coordinates of the nine modes of output k in the ligand's own atom order."""
    ser, xyz = FX[f"serial_{k}"], FX[f"xyz_{k}"].astype(np.float64) / 1000.0
    pos = {int(s): i for i, s in enumerate(ser)}
    order = [pos[s] for s in lig.serials()]
    return [m[order] for m in xyz]


def test_fixture_covers_every_output():
    assert len(CASES) == 31
    assert sorted({r for r, _, _ in CASES}) == ["1u1b", "2bm2", "2er7"]
    for _, _, k in CASES:
        assert FX[f"xyz_{k}"].shape[0] == 9 and FX[f"remark_{k}"].shape == (9, 3)


@pytest.mark.parametrize("rec,name,k", CASES, ids=[f"{r}-{l}" for r, l, _ in CASES])
def test_rmsd_bounds_reproduce(rec, name, k):
    lig = orc.Ligand(ligand_text(name))
    X = modes(k, lig)
    for x, (_, lb_ref, ub_ref) in zip(X, FX[f"remark_{k}"]):
        lb, ub = lig.rmsd_bounds(x, X[0])
        assert lb == pytest.approx(lb_ref, abs=2e-3) and ub == pytest.approx(ub_ref, abs=2e-3)


@pytest.mark.parametrize("rec,name,k", CASES, ids=[f"{r}-{l}" for r, l, _ in CASES])
def test_scores_reproduce(rec, name, k):
    g = grid(rec)
    lig = orc.Ligand(ligand_text(name))
    X = modes(k, lig)
    inter = [orc.noncache_eval(g, lig, x) for x in X]
    intra = [orc.intra_fast(g, lig, x) for x in X]
    best_intra = intra[int(np.argmin(np.add(inter, intra)))]     # out_cont[0] after the first sort
    for i, (a, b, (score, _, _)) in enumerate(zip(inter, intra, FX[f"remark_{k}"])):
        s = lig.conf_independent(a + b - best_intra)
        assert abs(s - score) <= 0.05 + 0.01, (i, s, score)


def test_tie_case_needs_the_sort_order_rule():
    """DB04226 on 2bm2 with the offset taken from the first printed mode misses its printed scores
    by 0.12-0.2: the rule above is what makes it reproduce, not the tolerance."""
    k = [kk for r, l, kk in CASES if (r, l) == ("2bm2", "DB04226")][0]
    g, lig = grid("2bm2"), orc.Ligand(ligand_text("DB04226"))
    X = modes(k, lig)
    s0 = lig.conf_independent(orc.noncache_eval(g, lig, X[0]))
    assert abs(s0 - FX[f"remark_{k}"][0, 0]) > 0.1
