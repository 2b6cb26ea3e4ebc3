"""More of the reference's own outputs as anchors for the oracle's final scoring (CPU only): the
QuickVina-W example's written results (QVW/input_file_example/2bm2_ligand_out.pdbqt with the
2bm2 receptor and box of 2bm2_config.txt, and 1u1b_ligand_out.pdbqt on the 1u1b scene; fixtures
tests/golden/qvw_*_ligand_out.pdbqt).  Each holds nine modes of a search independent of the V11
1u1b output, with the REMARK VINA RESULT score and rmsd bounds the reference computed
(QVW/main/main.cpp, the same eval_adjusted / conf_independent / rmsd bounds as V11/main/main.cpp:
440-470).  Rescoring their coordinates must give back those numbers: scores within the 0.05 of
the printed rounding, rmsd bounds within the coordinate rounding."""
import numpy as np
import pytest

import oracle as orc
from conftest import golden
from test_oracle_rescore import CENTER, SIZE, parse_models

CASES = {
    # QVW/input_file_example/2bm2_config.txt
    "2bm2": ("2bm2_protein.pdbqt", "2bm2_ligand.pdbqt", "qvw_2bm2_ligand_out.pdbqt", [40.415, 110.986, 82.673], [30.0, 30.0, 30.0]),
    # QVW/input_file_example/1u1b_config.txt (the V11 example's box)
    "1u1b": ("1u1b_protein.pdbqt", "1u1b_ligand.pdbqt", "qvw_1u1b_ligand_out.pdbqt", CENTER, SIZE),
}


def scene(name):
    prot, lig_f, out_f, center, size = CASES[name]
    xyz, xs = orc.parse_receptor(golden(prot))[:2]
    b, e, n = orc.grid_dims(center, size)
    g = orc.Grid(orc.Lut(), b, e, n)
    g.set_receptor(xyz, xs)
    lig = orc.Ligand(golden(lig_f))
    models, remarks = parse_models(golden(out_f))
    ser = lig.serials()
    X = [np.array([m[s] for s in ser]) for m in models]
    return g, lig, X, remarks


@pytest.mark.parametrize("name", sorted(CASES))
def test_rescoring_reproduces_quickvina_w_outputs(name):
    g, lig, X, remarks = scene(name)
    assert len(X) == 9
    best_intra = orc.intra_fast(g, lig, X[0])
    for x, (score, lb_ref, ub_ref) in zip(X, remarks):
        assert orc.within(g, lig, x)
        e = orc.noncache_eval(g, lig, x) + orc.intra_fast(g, lig, x) - best_intra
        assert abs(lig.conf_independent(e) - score) <= 0.05 + 1e-3, (lig.conf_independent(e), score)
        lb, ub = lig.rmsd_bounds(x, X[0])
        assert lb == pytest.approx(lb_ref, abs=2e-3) and ub == pytest.approx(ub_ref, abs=2e-3)
