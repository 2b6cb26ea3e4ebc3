"""Host side of the step after the search (CPU only): the PDBQT writer, the mode selection with
the final score, num_tors, and the argument checks of the post-processing entry points.

Anchored on the reference's own docking output V11/docking_result/1u1b_docking_result.pdbqt:
writing its nine modes back must give the file byte for byte, and selecting modes over its
poses (energies from the oracle at its coordinates) must give its printed scores and rmsd bounds.
"""
import numpy as np
import pytest

import oracle as orc
from conftest import golden
from test_oracle_rescore import CENTER, SIZE, parse_models


@pytest.fixture(scope="module")
def ref1u1b():
    lig = orc.Ligand(golden("1u1b_ligand.pdbqt"))
    models, remarks = parse_models(golden("1u1b_docking_result.pdbqt"))
    ser = lig.serials()
    X = np.array([[m[s] for s in ser] for m in models])
    lut = orc.Lut()
    xyz, xs = orc.parse_receptor(golden("1u1b_protein.pdbqt"))[:2]
    b, e, n = orc.grid_dims(CENTER, SIZE)
    g = orc.Grid(lut, b, e, n)
    g.set_receptor(xyz, xs)
    return lig, g, X, remarks


def test_writer_reproduces_reference_output_byte_for_byte(host_ctx, ref1u1b):
    """write_all_output + vina_remark + coords_to_pdbqt_string (V11/main/main.cpp:89-102, 158-170;
    V11/lib/model.cpp:1063-1110) on the reference's modes give back its output file."""
    import vdock
    lig, _, X, remarks = ref1u1b
    lid = host_ctx.add_ligand(golden("1u1b_ligand.pdbqt"))
    modes = [vdock.Mode(e, lb, ub, i) for i, (e, lb, ub) in enumerate(remarks)]
    text = host_ctx.write_pdbqt(lid, modes, X.astype(np.float32))
    assert text == golden("1u1b_docking_result.pdbqt")
    one = host_ctx.write_pdbqt(lid, modes[:1], X[:1].astype(np.float32))
    assert one.count("MODEL ") == 1 and one.endswith("ENDMDL\n")
    assert host_ctx.write_pdbqt(lid, [], X[:0].astype(np.float32)) == ""


def test_writer_rejects_coordinates_beyond_the_field(host_ctx, ref1u1b):
    import vdock
    lig, _, X, _ = ref1u1b
    lid = host_ctx.add_ligand(golden("1u1b_ligand.pdbqt"))
    bad = X[:1].astype(np.float32).copy()
    bad[0, 3, 0] = 12345.0          # "12345.000" needs 9 columns (string_write_coord's check)
    with pytest.raises(vdock.VdockError):
        host_ctx.write_pdbqt(lid, [vdock.Mode(-1.0, 0.0, 0.0, 0)], bad)


def test_num_tors_matches_oracle(host_ctx):
    import vdock
    texts = [golden(f"{n}_ligand.pdbqt") for n in ("1u1b", "2bm2", "2er7")]
    texts += [vdock.synth_ligand(12 + 5 * k, 3 + k, 77 + k) for k in range(8)]
    for t in texts:
        lid = host_ctx.add_ligand(t)
        assert host_ctx.ligand_info(lid).num_tors == pytest.approx(orc.Ligand(t).num_tors(), abs=0)
    assert host_ctx.ligand_info(host_ctx.add_ligand(texts[0])).num_tors == 15.0


def test_mode_selection_reproduces_reference_modes(host_ctx, ref1u1b):
    """main.cpp:425-470 over the reference's nine refined poses, shuffled, with the oracle's
    energies at their coordinates: the printed scores (as rounded) in order and the printed rmsd
    bounds."""
    import vdock
    lig, g, X, remarks = ref1u1b
    lid = host_ctx.add_ligand(golden("1u1b_ligand.pdbqt"))
    perm = np.random.default_rng(3).permutation(len(X))
    refined, coords = [], []
    for i in perm:
        ei = orc.noncache_eval(g, lig, X[i])
        ea = orc.intra_fast(g, lig, X[i])
        refined.append(vdock.Refined(ei + ea, ei, ea, 1))
        coords.append(X[i])
    modes = host_ctx.select_modes(lid, refined, np.array(coords, dtype=np.float32))
    assert len(modes) == 9
    assert [round(m.e, 1) for m in modes] == [r[0] for r in remarks]
    for m, r in zip(modes, remarks):
        assert abs(m.rmsd_lb - r[1]) < 2e-3 and abs(m.rmsd_ub - r[2]) < 2e-3
    assert [int(perm[m.pose]) for m in modes][0] == 0


def test_mode_selection_rules(host_ctx):
    """Poses that never reached the box are dropped; near-duplicates (rmsd < min_rmsd) keep the
    better one; energy_range and num_modes cut the list; a lone pose has zero rmsd bounds."""
    import vdock
    text = vdock.synth_ligand(14, 3, 5)
    lid = host_ctx.add_ligand(text)
    inf = host_ctx.ligand_info(lid)
    base = orc.Ligand(text).set_conf(orc.Ligand(text).identity_conf()).astype(np.float32)
    shifts = [0.0, 0.3, 5.0, 9.0, 14.0, 20.0]
    energies = [-6.0, -6.5, -5.0, -2.0, -4.0, 3.4e38]
    coords = np.array([base + np.array([s, 0, 0], dtype=np.float32) for s in shifts])
    ref = [vdock.Refined(e, e, 0.0, int(e < 1e30)) for e in energies]
    p = vdock.post_params(energy_range=3.0, num_modes=9, min_rmsd=1.0)
    modes = host_ctx.select_modes(lid, ref, coords, p)
    # pose 1 replaces its near-duplicate pose 0; pose 3 (-2.0) is beyond the range; pose 5 is out
    assert [m.pose for m in modes] == [1, 2, 4]
    assert modes[0].rmsd_lb == 0 and modes[0].rmsd_ub == 0
    d = shifts[2] - shifts[1]
    assert modes[1].rmsd_ub == pytest.approx(d, abs=1e-4) and modes[1].rmsd_lb <= modes[1].rmsd_ub
    w = 1 + vdock.DEFAULT_WEIGHTS[5] * inf.num_tors
    assert modes[1].e == pytest.approx(-5.0 / w, rel=1e-6)
    p.num_modes = 2
    assert [m.pose for m in host_ctx.select_modes(lid, ref, coords, p)] == [1, 2]
    one = host_ctx.select_modes(lid, ref[2:3], coords[2:3], p)
    assert len(one) == 1 and one[0].rmsd_lb == 0 and one[0].rmsd_ub == 0


def test_post_entry_points_need_a_device_and_valid_arguments(host_ctx):
    import vdock
    lid = host_ctx.add_ligand(vdock.synth_ligand(10, 2, 1))
    c = np.zeros((1, 7 + host_ctx.ligand_info(lid).n_torsions), dtype=np.float32)
    c[0, 3] = 1
    with pytest.raises(vdock.VdockError) as e:
        host_ctx.refine(lid, c)
    assert e.value.code == vdock.VD_ERR_NO_DEVICE
    with pytest.raises(vdock.VdockError) as e:
        host_ctx.refine(lid, c, vdock.post_params(num_modes=0))
    assert e.value.code == vdock.VD_ERR_ARG
    with pytest.raises(vdock.VdockError) as e:
        host_ctx.select_modes(lid, [vdock.Refined(-1, -1, 0, 1)], np.zeros((1, 10, 3), np.float32),
                              vdock.post_params(min_rmsd=-1.0))
    assert e.value.code == vdock.VD_ERR_ARG
    with pytest.raises(vdock.VdockError):
        host_ctx.post_results(0)
