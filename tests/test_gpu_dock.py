"""Statistical docking parity (GPU MC vs the FP64 oracle's MC with the same schedule, heuristics and
random stream).  Docking is stochastic and FP32 vs FP64 trajectories separate after a few steps,
so parity is judged over 16 seeds:
  - best score over the 16 runs within 0.1 kcal/mol of the oracle's best over its 16 runs;
  - mean per-run best score within 0.1 kcal/mol;
  - each side's overall top pose within 2 A RMSD (heavy atoms) of a pose where some chain of the
    other side ended, at most 0.1 kcal/mol above it: each search reaches the other's best minimum
    (distinct minima whose scores differ by less than the search's spread are equally good
    answers, and the argmin alone flips between them);
  - the fraction of runs whose top pose lands within 2 A of the overall top pose is not lower than
    the oracle's own fraction at the 1 % level (Fisher's exact test, one-sided; equal search
    accuracy).  With 16 runs a side the two fractions differ by two standard deviations about one
    time in twenty, so a fixed margin would fail by chance; the trajectories themselves change
    with the random stream, which is keyed on the ligand id and so on the order of the tests.
"""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def scene(ctx):
    """The reference's 1u1b example: its receptor and its search box (1u1b_config.txt)."""
    import vdock
    from conftest import golden
    ctx.set_receptor_pdbqt(golden("1u1b_protein.pdbqt"))
    box = vdock.box_from_center([-30.661, -29.540, 17.010], [24.0, 26.25, 22.5], 0.375)
    ctx.grid_populate(box)
    maps = ctx.grid_download()
    g = orc.Grid(orc.Lut(), *box.as_tuple())
    g.set_maps(maps.astype(np.float64))                    # the oracle docks on the same maps
    return dict(box=box, maps=maps, grid=g)


def no_worse_success(succ_g, succ_o, n=16, alpha=0.01):
    """One-sided Fisher's exact test: is the GPU's success count significantly below the oracle's?"""
    from scipy.stats import fisher_exact
    kg, ko = int(round(succ_g * n)), int(round(succ_o * n))
    p = fisher_exact([[kg, n - kg], [ko, n - ko]], alternative="less").pvalue
    say("success", kg, "of", n, "vs oracle", ko, "of", n, "one-sided p", p)
    return p >= alpha


def rmsd(a, b):
    return float(np.sqrt(((a - b) ** 2).sum(axis=-1).mean()))


def minima_distance(gb, gp, g_all, ob, op, o_all, de=0.1):
    """Largest of: the distance from the GPU's top pose to the nearest pose where some chain of
    some oracle run ended with an energy at most de above it, and the same with the sides
    swapped.  g_all / o_all = (energies, heavy-atom coordinates) of every chain of every run.
    Small ligands have several minima within a few hundredths of a kcal/mol, and which of them
    a side's argmin lands on is close to a coin toss; what equal searches share is that each
    reaches the other's best minimum."""
    gb, ob = np.asarray(gb), np.asarray(ob)

    def reached(x, e, side):
        energies, poses = side
        near = poses[energies <= e + de]
        return float(np.sqrt(((near - x) ** 2).sum(-1).mean(-1)).min()) if len(near) else np.inf

    i, j = int(np.argmin(gb)), int(np.argmin(ob))
    return max(reached(gp[i], gb[i], o_all), reached(op[j], ob[j], g_all))


@pytest.mark.parametrize("heavy,tors,lseed", [(12, 2, 31), (16, 4, 32)])
def test_docking_statistical_parity(ctx, scene, heavy, tors, lseed):
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    text = vdock.synth_ligand(heavy, tors, lseed)
    lid = ctx.add_ligand(text)
    lig = orc.Ligand(text)
    chains, depth, bfgs = 512, 8, 16
    b, e, _ = scene["box"].as_tuple()
    gb, ob, gp, op_, ga, oa = [], [], [], [], [], []
    for seed in range(16):
        p = vdock.default_params(chains=chains, search_depth=depth, bfgs_steps=bfgs, seed=seed)
        eg, cg, hg, _ = ctx.dock(lid, p)
        prm = orc.mc_params(chains, depth, bfgs, b, e, seed=seed, stream=lid, threads=8)
        eo, co, ho, _ = orc.mc_run(scene["grid"], lig, prm)
        gb.append(float(eg.min()))
        ob.append(float(eo.min()))
        gp.append(hg[int(np.argmin(eg))].astype(np.float64))
        op_.append(ho[int(np.argmin(eo))])
        ga.append((eg.astype(np.float64), hg.astype(np.float64)))
        oa.append((np.asarray(eo, dtype=np.float64), np.asarray(ho, dtype=np.float64)))
        say(f"seed {seed}: gpu {gb[-1]:.3f} oracle {ob[-1]:.3f}")
    g_all = (np.concatenate([x[0] for x in ga]), np.concatenate([x[1] for x in ga]))
    o_all = (np.concatenate([x[0] for x in oa]), np.concatenate([x[1] for x in oa]))
    gb, ob = np.array(gb), np.array(ob)
    top_g, top_o = gp[int(np.argmin(gb))], op_[int(np.argmin(ob))]
    succ_g = np.mean([rmsd(x, top_o) <= 2.0 for x in gp])
    succ_o = np.mean([rmsd(x, top_o) <= 2.0 for x in op_])
    say("best gpu", gb.min(), "oracle", ob.min(), "mean", gb.mean(), ob.mean(), "top-pose rmsd", rmsd(top_g, top_o),
        "to the other side's chains", minima_distance(gb, gp, g_all, ob, op_, o_all),
        "runs within 2A of top: gpu", succ_g, "oracle", succ_o)
    assert abs(gb.min() - ob.min()) <= 0.1
    assert abs(gb.mean() - ob.mean()) <= 0.1
    assert minima_distance(gb, gp, g_all, ob, op_, o_all) <= 2.0
    assert no_worse_success(succ_g, succ_o)


def _dock_seeds(ctx, scene, lid, lig, chains, depth, bfgs, policy, seeds=16):
    import vdock
    b, e, _ = scene["box"].as_tuple()
    out = dict(gb=[], ob=[], gp=[], op=[], ga=[], oa=[], gev=0, oev=0)
    for seed in range(seeds):
        p = vdock.default_params(chains=chains, search_depth=depth, bfgs_steps=bfgs, seed=seed, policy=policy)
        eg, cg, hg, st = ctx.dock(lid, p)
        prm = orc.mc_params(chains, depth, bfgs, b, e, seed=seed, stream=lid, threads=8, policy=policy)
        eo, co, ho, ev = orc.mc_run(scene["grid"], lig, prm)
        out["gb"].append(float(eg.min())); out["ob"].append(float(eo.min()))
        out["gp"].append(hg[int(np.argmin(eg))].astype(np.float64)); out["op"].append(ho[int(np.argmin(eo))])
        out["ga"].append((eg.astype(np.float64), hg.astype(np.float64)))
        out["oa"].append((np.asarray(eo, dtype=np.float64), np.asarray(ho, dtype=np.float64)))
        out["gev"] += st.evals; out["oev"] += ev
    for k in ("ga", "oa"):
        out[k] = (np.concatenate([x[0] for x in out[k]]), np.concatenate([x[1] for x in out[k]]))
    return out


@pytest.mark.parametrize("heavy,tors,lseed", [(10, 0, 51), (12, 1, 52)])
def test_quickvina2_policy_parity(ctx, scene, heavy, tors, lseed):
    """QuickVina2 policy (visited-history gate on both local searches) against the oracle's, with
    a search long enough for the 10 (6+T)-entry histories to fill so that the gate skips local
    searches.  Same 16-seed criteria as the Vina policy, plus: the gate saves evaluations on the
    device in the same proportion as in the oracle (within 0.05 of the ratio to the Vina policy)."""
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    text = vdock.synth_ligand(heavy, tors, lseed)
    lid = ctx.add_ligand(text)
    lig = orc.Ligand(text)
    chains, depth, bfgs = 256, 128, 12
    qv = _dock_seeds(ctx, scene, lid, lig, chains, depth, bfgs, policy=1)
    vn = _dock_seeds(ctx, scene, lid, lig, chains, depth, bfgs, policy=0, seeds=4)   # evaluations per run without the gate
    r_g = (qv["gev"] / 16) / (vn["gev"] / 4)
    r_o = (qv["oev"] / 16) / (vn["oev"] / 4)
    gb, ob = np.array(qv["gb"]), np.array(qv["ob"])
    top_g, top_o = qv["gp"][int(np.argmin(gb))], qv["op"][int(np.argmin(ob))]
    succ_g = np.mean([rmsd(x, top_o) <= 2.0 for x in qv["gp"]])
    succ_o = np.mean([rmsd(x, top_o) <= 2.0 for x in qv["op"]])
    say("qv2 best gpu", gb.min(), "oracle", ob.min(), "mean", gb.mean(), ob.mean(), "top-pose rmsd",
        rmsd(top_g, top_o), "succ", succ_g, succ_o, "eval ratio to vina: gpu", r_g, "oracle", r_o)
    assert abs(gb.min() - ob.min()) <= 0.1
    assert abs(gb.mean() - ob.mean()) <= 0.1
    assert minima_distance(gb, qv["gp"], qv["ga"], ob, qv["op"], qv["oa"]) <= 2.0
    assert no_worse_success(succ_g, succ_o)
    assert r_o < 0.97                      # the gate is active in this setting
    assert abs(r_g - r_o) <= 0.05


@pytest.mark.parametrize("heavy,tors,lseed", [(10, 0, 61), (12, 1, 62)])
def test_quickvina_w_policy_parity(ctx, scene, heavy, tors, lseed):
    """QuickVina-W policy (G-check on the ligand's shared global buffer, then the chain's own
    5 (6+T)-entry buffer) against the oracle's, whose chains advance one MC step at a time as the
    device's resident chains do.  The shared buffer makes both runs depend on the order in which
    chains record points, so the comparison is statistical: the 16-seed criteria of the Vina
    policy, the gate active in the oracle, and the device saving evaluations in about the same
    proportion (within 0.08 of the ratio to the Vina policy)."""
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    text = vdock.synth_ligand(heavy, tors, lseed)
    lid = ctx.add_ligand(text)
    lig = orc.Ligand(text)
    chains, depth, bfgs = 256, 96, 12
    qw = _dock_seeds(ctx, scene, lid, lig, chains, depth, bfgs, policy=2)
    vn = _dock_seeds(ctx, scene, lid, lig, chains, depth, bfgs, policy=0, seeds=4)
    r_g = (qw["gev"] / 16) / (vn["gev"] / 4)
    r_o = (qw["oev"] / 16) / (vn["oev"] / 4)
    gb, ob = np.array(qw["gb"]), np.array(qw["ob"])
    top_o = qw["op"][int(np.argmin(ob))]
    succ_g = np.mean([rmsd(x, top_o) <= 2.0 for x in qw["gp"]])
    succ_o = np.mean([rmsd(x, top_o) <= 2.0 for x in qw["op"]])
    say("qvw best gpu", gb.min(), "oracle", ob.min(), "mean", gb.mean(), ob.mean(), "top-pose distance",
        minima_distance(gb, qw["gp"], qw["ga"], ob, qw["op"], qw["oa"]), "succ", succ_g, succ_o, "eval ratio to vina: gpu", r_g,
        "oracle", r_o)
    assert abs(gb.min() - ob.min()) <= 0.1
    assert abs(gb.mean() - ob.mean()) <= 0.1
    assert minima_distance(gb, qw["gp"], qw["ga"], ob, qw["op"], qw["oa"]) <= 2.0
    assert no_worse_success(succ_g, succ_o)
    assert r_o < 0.97
    assert abs(r_g - r_o) <= 0.08


def test_initial_conformations_match_oracle(ctx, scene):
    """With zero MC steps the device keeps each chain's initial conformation: it must equal the
    oracle's draw from the shared Philox stream (FP32 rounding aside)."""
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    text = vdock.synth_ligand(14, 3, 77)
    lid = ctx.add_ligand(text)
    lig = orc.Ligand(text)
    b, e, _ = scene["box"].as_tuple()
    # one step whose local search cannot move: bfgs_steps=1 and compare the energies instead
    p = vdock.default_params(chains=32, search_depth=1, bfgs_steps=1, seed=3)
    eg, cg, _, _ = ctx.dock(lid, p)
    op = orc.mc_params(32, 1, 1, b, e, seed=3, stream=lid)
    eo, co, _, _ = orc.mc_run(scene["grid"], lig, op)
    close = np.abs(eg - eo) <= np.maximum(1e-2, 1e-3 * np.abs(eo))
    say("one-step energies within tolerance:", close.mean())
    assert close.mean() >= 0.9
