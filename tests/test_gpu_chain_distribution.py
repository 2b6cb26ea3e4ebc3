"""Docking parity at the level of single chains: the distribution of every chain's final energy.

The 16-seed criteria of the other docking tests judge each side by its best few chains, so their
statistics rest on the extremes of 16 runs; when a change re-rolls the trajectories (a different
FP32 rounding anywhere in the maps or the search), a fixed 0.1 kcal/mol margin on those extremes
fails about one time in twenty.  Here each side docks the same ligand over 32 seeds x 512 chains,
and the two samples of 16,384 final energies are compared as distributions:

  - two-sample Kolmogorov-Smirnov test, p >= 1e-3 (the critical distance for two samples of
    16,384 is 0.022 there; near the median the chains' distribution function rises about 0.66 per
    kcal/mol, so a shift of 0.03 kcal/mol in the body of the distribution fails it);
  - the 1, 5, 25 and 50 % quantiles within 0.05 kcal/mol of each other;
  - the mean per-run best within 3.5 standard errors of the difference;
  - the evaluation counts within 1 %: the same amount of search was done on both sides.

tests/test_oracle_distribution_power.py shows on the CPU, oracle against oracle, that these
criteria pass for two independent sets of seeds and fail for plausible planted mistakes.

The device runs FP32 on its own maps, the oracle FP64 on the same maps (uploaded as FP64), with
the same schedule, heuristics and random stream (the trajectories part after a few steps, so
agreement is in distribution only).  Scenes: the CPU-runnable configuration (2,000-atom synthetic
receptor, 55^3 box, a 20-atom, 5-torsion ligand at the reference's heuristic depth and BFGS steps)
through the plain launch, the same with the dense-BFGS local search (the reference's
--rilc_bfgs 0), a 16-atom, 4-torsion ligand in the reference's 1u1b pocket through the batched
persistent launch, and the QuickVina2 policy in that pocket deep enough for its visited histories
to fill.  Measured on MI355X: Vina policy KS p 0.90-0.98 (64 and 32 seeds), quantiles within
0.01, evaluation counts within 0.25 %; dense BFGS KS p 0.22; QuickVina2 KS p 0.20 with evaluation
counts within 0.001 % (the gate skips the same share of searches on both sides).
"""
import numpy as np
import pytest

import oracle as orc
from helpers import compare_chain_distributions as compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(module_ctx):
    """This module's own context (conftest.module_ctx): its random streams do not depend on the
    tests that ran before it."""
    return module_ctx

SEEDS, CHAINS = 32, 512


def test_cpu_runnable_config_chain_distribution(ctx):
    import vdock
    xyz, xs = vdock.synth_receptor(2000, 36.0, 8.0, 42)
    ctx.set_receptor(xyz, xs)
    box = vdock.box_from_center([0, 0, 0], [20.0] * 3, 0.375)
    ctx.grid_populate(box)
    g = orc.Grid(orc.Lut(), *box.as_tuple())
    g.set_maps(ctx.grid_download().astype(np.float64))
    text = vdock.synth_ligand(20, 5, 1)
    lid = ctx.add_ligand(text)
    lig = orc.Ligand(text)
    inf = ctx.ligand_info(lid)
    depth, bfgs = vdock.heuristics(inf.n_atoms, inf.n_torsions)
    b, e, _ = box.as_tuple()
    G, O, gb, ob, gev, oev = [], [], [], [], 0, 0
    for seed in range(SEEDS):
        eg, _, _, st = ctx.dock(lid, vdock.default_params(chains=CHAINS, seed=seed))
        eo, _, _, ev = orc.mc_run(g, lig, orc.mc_params(CHAINS, depth, bfgs, b, e, seed=seed, stream=lid, threads=8))
        G.append(np.asarray(eg, np.float64)); O.append(np.asarray(eo, np.float64))
        gb.append(G[-1].min()); ob.append(O[-1].min())
        gev += st.evals; oev += ev
    compare("cpu-runnable configuration, plain launch", G, O, gb, ob, gev, oev)


def test_batched_launch_chain_distribution(ctx):
    """The batched persistent launch (the screening path): the ligand docks in a batch with two
    others of another lanes-per-chain class, so that both classes' launches run at the same time."""
    import vdock
    from conftest import golden
    ctx.set_receptor_pdbqt(golden("1u1b_protein.pdbqt"))
    box = vdock.box_from_center([-30.661, -29.540, 17.010], [24.0, 26.25, 22.5], 0.375)
    ctx.grid_populate(box)
    g = orc.Grid(orc.Lut(), *box.as_tuple())
    g.set_maps(ctx.grid_download().astype(np.float64))
    text = vdock.synth_ligand(16, 4, 32)
    lid = ctx.add_ligand(text)
    others = [ctx.add_ligand(vdock.synth_ligand(40 + k, 12, 700 + k)) for k in range(2)]
    lig = orc.Ligand(text)
    depth, bfgs = 8, 16
    b, e, _ = box.as_tuple()
    G, O, gb, ob, gev, oev = [], [], [], [], 0, 0
    for seed in range(SEEDS):
        p = vdock.default_params(chains=CHAINS, search_depth=depth, bfgs_steps=bfgs, seed=seed)
        ctx.dock_batch([lid] + others, p)
        # the 16/4 ligand is the only one of the 16-lane class: that launch's count is its own
        mine = [r for r in ctx.last_launches() if r.lanes_per_chain == 16]
        assert len(ctx.last_launches()) == 2 and len(mine) == 1 and mine[0].ligands == 1
        eg, _, _ = ctx.batch_results(0)
        eo, _, _, ev = orc.mc_run(g, lig, orc.mc_params(CHAINS, depth, bfgs, b, e, seed=seed, stream=lid, threads=8))
        G.append(np.asarray(eg, np.float64)); O.append(np.asarray(eo, np.float64))
        gb.append(G[-1].min()); ob.append(O[-1].min())
        gev += mine[0].evals; oev += ev
    compare("1u1b pocket, 16/4 ligand, batched launch", G, O, gb, ob, gev, oev)



def _config0_scene(ctx):
    import vdock
    xyz, xs = vdock.synth_receptor(2000, 36.0, 8.0, 42)
    ctx.set_receptor(xyz, xs)
    box = vdock.box_from_center([0, 0, 0], [20.0] * 3, 0.375)
    ctx.grid_populate(box)
    g = orc.Grid(orc.Lut(), *box.as_tuple())
    g.set_maps(ctx.grid_download().astype(np.float64))
    return box, g


def test_dense_bfgs_chain_distribution(ctx):
    """The dense-BFGS local search (the reference's --rilc_bfgs 0) with its own search-depth and
    step heuristics, on the CPU-runnable configuration."""
    import vdock
    box, g = _config0_scene(ctx)
    text = vdock.synth_ligand(20, 5, 1)
    lid = ctx.add_ligand(text)
    lig = orc.Ligand(text)
    depth, bfgs = vdock.heuristics(lig.n_atoms, lig.T, rilc=False)
    b, e, _ = box.as_tuple()
    G, O, gb, ob, gev, oev = [], [], [], [], 0, 0
    for seed in range(SEEDS):
        eg, _, _, st = ctx.dock(lid, vdock.default_params(chains=CHAINS, seed=seed, local_search=vdock.VD_LS_BFGS))
        eo, _, _, ev = orc.mc_run(g, lig, orc.mc_params(CHAINS, depth, bfgs, b, e, seed=seed, stream=lid, threads=8,
                                                        mode=0))
        G.append(np.asarray(eg, np.float64)); O.append(np.asarray(eo, np.float64))
        gb.append(G[-1].min()); ob.append(O[-1].min())
        gev += st.evals; oev += ev
    compare("dense BFGS, cpu-runnable configuration", G, O, gb, ob, gev, oev)


def test_quickvina2_chain_distribution(ctx):
    """The QuickVina2 policy deep enough for the visited histories to fill (10 (6+T) entries,
    reached after 40 of a chain's 128 local searches), so that the gate skips searches on both
    sides: 1u1b pocket, a 12-atom, 2-torsion ligand, depth 64, 256 chains x 32 seeds."""
    import vdock
    from conftest import golden
    ctx.set_receptor_pdbqt(golden("1u1b_protein.pdbqt"))
    box = vdock.box_from_center([-30.661, -29.540, 17.010], [24.0, 26.25, 22.5], 0.375)
    ctx.grid_populate(box)
    g = orc.Grid(orc.Lut(), *box.as_tuple())
    g.set_maps(ctx.grid_download().astype(np.float64))
    text = vdock.synth_ligand(12, 2, 61)
    lid = ctx.add_ligand(text)
    lig = orc.Ligand(text)
    chains, depth, bfgs = 256, 64, 12
    b, e, _ = box.as_tuple()
    G, O, gb, ob, gev, oev = [], [], [], [], 0, 0
    for seed in range(SEEDS):
        p = vdock.default_params(chains=chains, search_depth=depth, bfgs_steps=bfgs, seed=seed, policy=1)
        eg, _, _, st = ctx.dock(lid, p)
        eo, _, _, ev = orc.mc_run(g, lig, orc.mc_params(chains, depth, bfgs, b, e, seed=seed, stream=lid, threads=8,
                                                        policy=1))
        G.append(np.asarray(eg, np.float64)); O.append(np.asarray(eo, np.float64))
        gb.append(G[-1].min()); ob.append(O[-1].min())
        gev += st.evals; oev += ev
    compare("QuickVina2 policy, 1u1b pocket", G, O, gb, ob, gev, oev)
