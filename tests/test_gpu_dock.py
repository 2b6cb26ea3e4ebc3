"""Statistical docking parity (GPU MC vs the FP64 oracle's MC with the same schedule, heuristics and
random stream).  Docking is stochastic and FP32 vs FP64 trajectories separate after a few steps,
so parity is judged over 16 seeds with helpers.multi_ligand_parity: the runs' best scores pass a
two-sided Mann-Whitney test; the mean per-run best score and the best score over the 16 runs are
within 0.1 kcal/mol of the oracle's beyond three bootstrap standard errors; the runs whose top pose
lies within 2 A of the best pose either side found are not significantly fewer on the device
(Fisher, one-sided, 1 %).

Fixed 0.1 kcal/mol margins on the best of 16 runs, with no allowance for their scatter, fail by
chance whenever a change anywhere re-rolls the trajectories (a different FP32 rounding of the
maps, for one): a rare deep minimum is then found by one side's 16 runs and not by the other's.
test_gpu_chain_distribution.py compares the two searches chain by chain over 32 seeds, which is
the powerful check.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import assert_parity, multi_ligand_parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(module_ctx):
    """This module's own context (conftest.module_ctx): its random streams do not depend on the
    tests that ran before it."""
    return module_ctx


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


@pytest.mark.parametrize("heavy,tors,lseed", [(12, 2, 31), (16, 4, 32)])
def test_docking_statistical_parity(ctx, scene, heavy, tors, lseed):
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    text = vdock.synth_ligand(heavy, tors, lseed)
    lid = ctx.add_ligand(text)
    lig = orc.Ligand(text)
    chains, depth, bfgs = 512, 8, 16
    b, e, _ = scene["box"].as_tuple()
    gb, ob, gp, op_ = [], [], [], []
    for seed in range(16):
        p = vdock.default_params(chains=chains, search_depth=depth, bfgs_steps=bfgs, seed=seed)
        eg, cg, hg, _ = ctx.dock(lid, p)
        prm = orc.mc_params(chains, depth, bfgs, b, e, seed=seed, stream=lid, threads=8)
        eo, co, ho, _ = orc.mc_run(scene["grid"], lig, prm)
        gb.append(float(eg.min()))
        ob.append(float(eo.min()))
        gp.append(hg[int(np.argmin(eg))].astype(np.float64))
        op_.append(ho[int(np.argmin(eo))])
        say(f"seed {seed}: gpu {gb[-1]:.3f} oracle {ob[-1]:.3f}")
    assert_parity(multi_ligand_parity([list(zip(gb, gp))], [list(zip(ob, op_))], ["%d/%d" % (lig.n_atoms, lig.T)]))


def _dock_seeds(ctx, scene, lid, lig, chains, depth, bfgs, policy, seeds=16):
    import vdock
    b, e, _ = scene["box"].as_tuple()
    out = dict(gb=[], ob=[], gp=[], op=[], gev=0, oev=0)
    for seed in range(seeds):
        p = vdock.default_params(chains=chains, search_depth=depth, bfgs_steps=bfgs, seed=seed, policy=policy)
        eg, cg, hg, st = ctx.dock(lid, p)
        prm = orc.mc_params(chains, depth, bfgs, b, e, seed=seed, stream=lid, threads=8, policy=policy)
        eo, co, ho, ev = orc.mc_run(scene["grid"], lig, prm)
        out["gb"].append(float(eg.min())); out["ob"].append(float(eo.min()))
        out["gp"].append(hg[int(np.argmin(eg))].astype(np.float64)); out["op"].append(ho[int(np.argmin(eo))])
        out["gev"] += st.evals; out["oev"] += ev
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
    say("qv2 eval ratio to vina: gpu", r_g, "oracle", r_o)
    assert_parity(multi_ligand_parity([list(zip(qv["gb"], qv["gp"]))], [list(zip(qv["ob"], qv["op"]))],
                                      ["%d/%d" % (lig.n_atoms, lig.T)]))
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
    say("qvw eval ratio to vina: gpu", r_g, "oracle", r_o)
    assert_parity(multi_ligand_parity([list(zip(qw["gb"], qw["gp"]))], [list(zip(qw["ob"], qw["op"]))],
                                      ["%d/%d" % (lig.n_atoms, lig.T)]))
    assert r_o < 0.97
    assert abs(r_g - r_o) <= 0.08


def test_quickvina_w_batched_parity(ctx, scene):
    """QuickVina-W through the batched launch, with both lanes-per-chain classes running at once
    (an 18-atom, 3-torsion ligand in the 16-lane class, a 36-atom, 11-torsion one in the 64-lane
    class, each with its own shared buffer), against the oracle's runs of the same ligands and
    streams: the 16-seed criteria of the Vina policy per ligand, and the gate saving evaluations in
    about the same proportion on both sides (ratio to the Vina policy within 0.08)."""
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    texts = [vdock.synth_ligand(18, 3, 71), vdock.synth_ligand(36, 11, 72)]
    ids = [ctx.add_ligand(t) for t in texts]
    ligs = [orc.Ligand(t) for t in texts]
    chains, depth, bfgs = 256, 48, 12
    b, e, _ = scene["box"].as_tuple()

    def run(policy, seeds):
        dev, ora = [[] for _ in ids], [[] for _ in ids]
        gev = oev = 0
        for seed in range(seeds):
            st = ctx.dock_batch(ids, vdock.default_params(chains=chains, search_depth=depth, bfgs_steps=bfgs,
                                                          seed=seed, policy=policy))
            assert sorted(r.kernel.split("<")[1].split(",")[0] for r in ctx.last_launches()) == ["16", "64"]
            gev += st.evals
            for k, (lid, lig) in enumerate(zip(ids, ligs)):
                eg, _, hg = ctx.batch_results(k)
                prm = orc.mc_params(chains, depth, bfgs, b, e, seed=seed, stream=lid, threads=8, policy=policy)
                eo, _, ho, ev = orc.mc_run(scene["grid"], lig, prm)
                oev += ev
                dev[k].append((float(eg.min()), hg[int(np.argmin(eg))].astype(np.float64)))
                ora[k].append((float(eo.min()), ho[int(np.argmin(eo))]))
        return dev, ora, gev, oev

    dev, ora, gq, oq = run(2, 16)
    _, _, gv, ov = run(0, 4)
    r_g, r_o = (gq / 16) / (gv / 4), (oq / 16) / (ov / 4)
    say("batched qvw eval ratio to vina: gpu", r_g, "oracle", r_o)
    assert_parity(multi_ligand_parity(dev, ora, ["%d/%d" % (l.n_atoms, l.T) for l in ligs]))
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
