"""Parity of the device refinement and final scores (vd_refine_kernel, FP64) with the FP64 oracle.

The oracle's refinement is pinned on the reference's own 1u1b output (test_oracle_rescore.py);
here the device runs the same poses through the C ABI.  Both sides compute in FP64 with the same
operation order except for summation orders, and the C ABI hands results back in FP32, so:
energies within 1e-4 kcal/mol and poses within 1e-3 A for poses near a minimum; for poses the
BFGS moves far, a knife-edge line-search decision can part the two runs, so the bound there is
on the distribution; final scores within 1e-3 and as the reference prints them (0.1 kcal/mol).
"""
import os

import numpy as np
import pytest

import oracle as orc
from conftest import golden
from test_oracle_rescore import CENTER, SIZE, conf_from_coords, parse_models

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def site(ctx):
    import vdock
    ctx.set_receptor_pdbqt(golden("1u1b_protein.pdbqt"))
    box = vdock.box_from_center(CENTER, SIZE, 0.375)
    ctx.grid_populate(box)
    lig = orc.Ligand(golden("1u1b_ligand.pdbqt"))
    xyz, xs = orc.parse_receptor(golden("1u1b_protein.pdbqt"))[:2]
    b, e, n = orc.grid_dims(CENTER, SIZE)
    g = orc.Grid(orc.Lut(), b, e, n)
    g.set_receptor(xyz, xs)
    models, remarks = parse_models(golden("1u1b_docking_result.pdbqt"))
    ser = lig.serials()
    X = [np.array([m[s] for s in ser]) for m in models]
    # the C ABI takes FP32 conformations: both sides start from the same rounded values
    confs = np.array([conf_from_coords(lig, x) for x in X]).astype(np.float32).astype(np.float64)
    lid = ctx.add_ligand(golden("1u1b_ligand.pdbqt"))
    return dict(box=box, lig=lig, grid=g, X=X, remarks=remarks, confs=confs, lid=lid)


def rmsd(a, b):
    return float(np.sqrt(((np.asarray(a, np.float64) - b) ** 2).sum(-1).mean()))


def test_refine_matches_oracle_on_reference_poses(ctx, site):
    lig, g, lid = site["lig"], site["grid"], site["lid"]
    steps = orc.refine_steps(lig)
    co, xyz, res = ctx.refine(lid, site["confs"])
    for i, c in enumerate(site["confs"]):
        eo, cfo, xo = orc.refine(g, lig, c, steps)
        r = res[i]
        say(f"pose {i}: gpu {r.e:.4f} oracle {eo:.4f} rmsd {rmsd(xyz[i], xo):.4f}")
        assert r.inside == 1 and orc.within(g, lig, xyz[i])
        assert abs(r.e - eo) <= 1e-4
        assert rmsd(xyz[i], xo) <= 1e-3
        assert np.allclose(lig.set_conf(co[i]), xyz[i], atol=1e-3)
        # the final score's parts at the refined pose
        ei, ea = orc.noncache_eval(g, lig, xo), orc.intra_fast(g, lig, xo)
        assert abs(r.e_inter - ei) <= 1e-4 + 1e-6 * abs(ei), (r.e_inter, ei)
        assert abs(r.e_intra - ea) <= 1e-4 + 1e-6 * abs(ea), (r.e_intra, ea)


def test_refine_perturbed_poses_statistical_parity(ctx, site):
    """Poses pushed off the minimum (torsions +- 0.25 rad, shift 0.5 A): long descents where a
    line-search decision on the edge can part the runs, so the bound is on the distribution."""
    lig, g, lid = site["lig"], site["grid"], site["lid"]
    rng = np.random.default_rng(11)
    confs = []
    for k in range(24):
        c = site["confs"][k % 9].copy()
        c[:3] += rng.normal(0, 0.5, 3)
        c[7:] += rng.normal(0, 0.25, lig.T)
        confs.append(c)
    confs = np.array(confs).astype(np.float32).astype(np.float64)
    co, xyz, res = ctx.refine(lid, confs)
    de, dr = [], []
    for i, c in enumerate(confs):
        eo, _, xo = orc.refine(g, lig, c, orc.refine_steps(lig))
        de.append(res[i].e - eo)
        dr.append(rmsd(xyz[i], xo))
    de, dr = np.abs(de), np.array(dr)
    say("perturbed: |de| median", np.median(de), "max", de.max(), "rmsd median", np.median(dr), "max", dr.max())
    assert np.median(de) <= 1e-4 and np.median(dr) <= 1e-3
    assert np.mean(de <= 1e-3) >= 0.9 and de.max() <= 0.1


def test_refine_brings_poses_into_the_box_or_reports_max(ctx, site):
    import vdock
    lig, g, lid = site["lig"], site["grid"], site["lid"]
    c = site["confs"][0].copy()
    c[0] += g.end[0] - lig.set_conf(c)[:, 0].max() + 1.0
    far = site["confs"][0].copy()
    far[0] += 80.0
    c, far = (v.astype(np.float32).astype(np.float64) for v in (c, far))
    co, xyz, res = ctx.refine(lid, np.array([c]))
    assert res[0].inside == 1 and res[0].e < 1e30 and orc.within(g, lig, xyz[0].astype(np.float64))
    eo, _, _ = orc.refine(g, lig, c, orc.refine_steps(lig))
    say("pulled in: gpu", res[0].e, "oracle", eo)
    assert abs(res[0].e - eo) <= 0.1
    co, xyz, res = ctx.refine(lid, np.array([far]), vdock.post_params(max_steps=1))
    assert res[0].inside == 0 and res[0].e > 3e38


def test_postprocess_reproduces_reference_modes(ctx, site):
    """Refinement + selection on the device path for the reference's nine poses: its printed
    scores (as rounded), best first, rmsd bounds close to the oracle's post-processing."""
    lig, g, lid = site["lig"], site["grid"], site["lid"]
    st = ctx.postprocess([lid], [site["confs"]])
    modes, conf, xyz = ctx.post_results(0)
    omodes = orc.postprocess(g, lig, site["confs"], orc.refine_steps(lig))
    say("gpu modes", [(round(m.e, 3), round(m.rmsd_lb, 3), round(m.rmsd_ub, 3)) for m in modes])
    say("oracle   ", [(round(m["e"], 3), round(m["lb"], 3), round(m["ub"], 3)) for m in omodes])
    assert st.blocks == 9 and st.evals > 0
    assert len(modes) == 9
    assert [round(m.e, 1) for m in modes] == [r[0] for r in site["remarks"]]
    for m, o in zip(modes, omodes):
        assert abs(m.e - o["e"]) <= 1e-3
        assert abs(m.rmsd_lb - o["lb"]) <= 1e-3 and abs(m.rmsd_ub - o["ub"]) <= 1e-3
    text = ctx.write_pdbqt(lid, modes, xyz)
    assert text.count("MODEL ") == 9 and text.count("REMARK VINA RESULT:") == 9


def test_dock_then_postprocess_1u1b(ctx, site):
    """The whole reference sequence on the device: dock (the reference example's 8192 chains),
    the host's pose list, refinement, final scores, output.  Same scores as the oracle's
    post-processing of the same poses; a best mode as good as the reference's printed -7.9 to
    within the search's own spread."""
    import vdock
    lig, g, lid = site["lig"], site["grid"], site["lid"]
    p = vdock.default_params(chains=8192, seed=7)
    ctx.dock_batch([lid], p)
    st = ctx.postprocess_docked(20)
    modes, conf, xyz = ctx.post_results(0)
    e, cf, hv = ctx.batch_results(0, p)
    kept = vdock.cluster(e, hv, 1.0, 20)
    omodes = orc.postprocess(g, lig, cf[kept].astype(np.float64), orc.refine_steps(lig))
    say("docked: gpu", [round(m.e, 2) for m in modes], "oracle", [round(m["e"], 2) for m in omodes],
        "refine ms", st.kernel_ms)
    assert len(modes) >= 1 and modes[0].e <= -7.0
    assert len(modes) == len(omodes)
    assert np.abs(np.array([m.e for m in modes]) - [o["e"] for o in omodes]).max() <= 1e-3
    assert all(a.e <= b.e for a, b in zip(modes, modes[1:]))
    assert modes[0].rmsd_lb == 0 and modes[0].rmsd_ub == 0
    text = ctx.write_pdbqt(lid, modes, xyz)
    assert text.startswith("MODEL 1\nREMARK VINA RESULT:") and text.count("ENDMDL") == len(modes)


def test_driver_runs_the_reference_example(ctx, site, tmp_path):
    """The reference's example run (1u1b_config.txt: its box and 8192 chains) through the
    command-line flow: the output file has the reference output's layout line for line (every
    line but the coordinates and the scores equal to V11/docking_result/1u1b_docking_result.pdbqt's
    first model), and a best mode as good as the printed -7.9 to within the search's spread.
    Directory mode writes one <name>_out.pdbqt per ligand into <dir>_out."""
    import shutil
    from conftest import GOLDEN
    from vdock import driver
    lig = str(tmp_path / "1u1b_ligand.pdbqt")
    shutil.copy(f"{GOLDEN}/1u1b_ligand.pdbqt", lig)
    opts = driver.resolve(f"{GOLDEN}/1u1b_config.txt", receptor=f"{GOLDEN}/1u1b_protein.pdbqt", ligand=lig,
                          out=str(tmp_path / "out.pdbqt"), seed=11)
    res = driver.run(opts, ctx=ctx, log=lambda *a: None)
    text = open(tmp_path / "out.pdbqt").read()
    assert text == res[0]["text"] and res[0]["modes"][0].e <= -7.0
    ref = golden("1u1b_docking_result.pdbqt").splitlines()
    ref_model = ref[:ref.index("ENDMDL") + 1]
    got = text.splitlines()
    got_model = got[:got.index("ENDMDL") + 1]
    assert len(got_model) == len(ref_model)
    for a, b in zip(got_model, ref_model):
        if a.startswith("REMARK VINA RESULT:"):
            assert b.startswith("REMARK VINA RESULT:") and len(a) == len(b)
        elif a.startswith(("ATOM", "HETATM")):
            assert a[:30] == b[:30] and a[54:] == b[54:]
        else:
            assert a == b
    d = tmp_path / "ligs"
    d.mkdir()
    shutil.copy(f"{GOLDEN}/1u1b_ligand.pdbqt", d / "a.pdbqt")
    shutil.copy(f"{GOLDEN}/2bm2_ligand.pdbqt", d / "b.pdbqt")
    (d / "bad.pdbqt").write_text("ROOT\nnot a pdbqt line\n")
    opts = driver.resolve(receptor=f"{GOLDEN}/1u1b_protein.pdbqt", ligand_directory=str(d), thread=2048,
                          center_x=-30.661, center_y=-29.540, center_z=17.010, size_x=24.0, size_y=26.25,
                          size_z=22.5, seed=3)
    res = driver.run(opts, ctx=ctx, log=lambda *a: None)
    assert [os.path.basename(r["out"]) for r in res] == ["a_out.pdbqt", "b_out.pdbqt"]
    assert all(os.path.exists(r["out"]) and r["modes"] for r in res)
    assert os.path.dirname(res[0]["out"]) == str(d) + "_out"


@pytest.mark.parametrize("heavy,tors,seed", [(75, 14, 5), (12, 0, 6), (40, 30, 7)])
def test_refine_matches_oracle_on_synthetic_ligands(ctx, heavy, tors, seed):
    """Shapes beyond the reference example: more atoms than a wavefront has lanes (pairs in the
    general chunk layout), a rigid ligand (no torsions), many torsions.  Docked poses of each,
    refined on the device and by the oracle from the same FP32 starts."""
    import vdock
    xyz, xs = vdock.synth_receptor(3000, 40.0, 8.0, 11)
    ctx.set_receptor(xyz, xs)
    box = vdock.box_from_center([0, 0, 0], [18.0, 18.0, 18.0], 0.375)
    ctx.grid_populate(box)
    text = vdock.synth_ligand(heavy, tors, seed)
    lid = ctx.add_ligand(text)
    lig = orc.Ligand(text)
    b, e, n = box.as_tuple()
    g = orc.Grid(orc.Lut(), b, e, n)
    g.set_receptor(xyz.astype(np.float64), xs)
    p = vdock.default_params(chains=64, search_depth=4, bfgs_steps=8, seed=seed)
    en, conf, _, _ = ctx.dock(lid, p)
    starts = conf[np.argsort(en)[:6]].astype(np.float64)
    co, out_xyz, res = ctx.refine(lid, starts)
    de, dr = [], []
    for i, c in enumerate(starts):
        eo, _, xo = orc.refine(g, lig, c, orc.refine_steps(lig))
        if res[i].inside:
            de.append(abs(res[i].e - eo))
            dr.append(rmsd(out_xyz[i], xo))
            ei, ea = orc.noncache_eval(g, lig, out_xyz[i].astype(np.float64)), orc.intra_fast(g, lig, out_xyz[i].astype(np.float64))
            assert abs(res[i].e_inter - ei) <= 1e-3 + 1e-5 * abs(ei) and abs(res[i].e_intra - ea) <= 1e-3 + 1e-5 * abs(ea)
        else:
            assert eo > 1e300
    say(f"synthetic {heavy}/{tors}: atoms {lig.n_atoms} |de| {np.round(de, 6)} rmsd {np.round(dr, 5)}")
    assert len(de) >= 4
    assert np.median(de) <= 1e-4 and np.median(dr) <= 1e-3 and max(de) <= 0.1
