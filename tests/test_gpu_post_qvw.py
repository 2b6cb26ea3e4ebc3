"""The step after the search on the device for the QuickVina-W example's written 2bm2 result
(QVW/input_file_example/2bm2_ligand_out.pdbqt, its 9,311-atom receptor and box): refinement of
its nine poses, final scores and mode selection agree with the oracle's post-processing of the
same poses to 1e-3 (scores and rmsd bounds).

The printed scores themselves are those of the written coordinates (rescored by the oracle in
test_oracle_rescore_qvw.py).  Refining them again is not the identity here: QuickVina-W's own
refinement stopped at its step limit on some poses, and a second pass takes mode 7 from -7.43 to
-7.51, so the rounded scores after it need not be the printed ones."""
import numpy as np
import pytest

import oracle as orc
from conftest import golden
from test_oracle_rescore import conf_from_coords, parse_models

pytestmark = pytest.mark.gpu

CENTER = [40.415, 110.986, 82.673]   # QVW/input_file_example/2bm2_config.txt
SIZE = [30.0, 30.0, 30.0]


def say(*a):
    print(*a, flush=True)


def test_postprocess_reproduces_quickvina_w_2bm2_modes(ctx):
    import vdock
    ctx.set_receptor_pdbqt(golden("2bm2_protein.pdbqt"))
    box = vdock.box_from_center(CENTER, SIZE, 0.375)
    say("2bm2 populate", box.dims)
    ctx.grid_populate(box)
    lig = orc.Ligand(golden("2bm2_ligand.pdbqt"))
    lid = ctx.add_ligand(golden("2bm2_ligand.pdbqt"))
    xyz, xs = orc.parse_receptor(golden("2bm2_protein.pdbqt"))[:2]
    g = orc.Grid(orc.Lut(), *box.as_tuple())
    g.set_receptor(xyz, xs)
    models, remarks = parse_models(golden("qvw_2bm2_ligand_out.pdbqt"))
    ser = lig.serials()
    X = [np.array([m[s] for s in ser]) for m in models]
    # the C ABI takes FP32 conformations: both sides start from the same rounded values
    confs = np.array([conf_from_coords(lig, x) for x in X]).astype(np.float32).astype(np.float64)
    say("2bm2 postprocess")
    st = ctx.postprocess([lid], [confs])
    modes, conf, out_xyz = ctx.post_results(0)
    omodes = orc.postprocess(g, lig, confs, orc.refine_steps(lig))
    say("gpu modes", [(round(m.e, 3), round(m.rmsd_lb, 3), round(m.rmsd_ub, 3)) for m in modes])
    say("printed  ", remarks)
    assert st.blocks == 9 and len(modes) == 9 and len(omodes) == 9
    assert modes[0].e <= remarks[0][0] + 0.05           # the best mode only improves
    for m, o in zip(modes, omodes):
        assert abs(m.e - o["e"]) <= 1e-3
        assert abs(m.rmsd_lb - o["lb"]) <= 1e-3 and abs(m.rmsd_ub - o["ub"]) <= 1e-3
