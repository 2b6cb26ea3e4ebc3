"""The large-box screening workload as a parity case: 20,000-atom synthetic receptor (90 A cube,
25 A cavity), 70 A box at 0.375 A (188^3 points, 17 FP32 maps = 452 MB), the QuickVina2 policy.

  - every point of the 17 device maps against the FP64 oracle's cache::populate:
    |d| <= 1e-4 max(1, |ref|) + slack (the oracle's bound on what an FP32 rounding of r^2 within
    0.003 LUT bins of a bin edge or the cutoff can change);
  - QuickVina2 parity on large ligands (44 heavy atoms / 12 torsions and 40 / 14), 128 chains,
    search depth 300 so that the chains' histories (10 (6+T) entries) fill and the gate skips
    local searches, 16 seeds against the oracle's QuickVina2 Monte Carlo on the same maps: the
    gate removes the same share of evaluations on both sides (ratio to the Vina policy within
    0.05), and the runs' best scores and top poses meet helpers.multi_ligand_parity (runs scatter
    by about 0.5 kcal/mol in this box, so 16 seeds fix a mean to about 0.13: the 0.1 kcal/mol
    bounds hold to within three bootstrap standard errors);
  - the 4,096-ligand batch (10-80 heavy atoms, 0-24 torsions drawn, 128 chains each, QuickVina2
    policy, one persistent launch per lanes-per-chain class): finite results for every ligand,
    and on a sample of ligands the best chains' energies are the oracle's evaluation of the
    returned conformations on the same maps and their heavy-atom coordinates those of the
    conformation.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import assert_parity, check_best_chains, maps_within, multi_ligand_parity, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(module_ctx):
    """This module's own context (conftest.module_ctx): its random streams do not depend on the
    tests that ran before it."""
    return module_ctx


@pytest.fixture(scope="module")
def large_box(ctx):
    import vdock
    from vdock import workloads as W
    say("large box: 20000-atom receptor, 188^3 maps")
    xyz, xs, box = W.large_box_scene(vdock, ctx)
    assert tuple(box.dims) == (188, 188, 188)
    maps = ctx.grid_download()
    shared = orc.Grid(orc.Lut(), *box.as_tuple())
    shared.set_maps(maps.astype(np.float64))
    return dict(xyz=xyz, xs=xs, box=box, maps=maps, shared=shared)


def test_large_box_maps_match_oracle(large_box):
    w = large_box
    g = orc.Grid(orc.Lut(), *w["box"].as_tuple())
    g.set_receptor(w["xyz"].astype(np.float64), w["xs"])
    say("oracle populate 188^3")
    slack = g.populate_with_slack(0.003)
    ref = g.maps()
    bad, worst, strict = maps_within(w["maps"], ref, slack)
    say("large box maps: points outside the bound", bad, "max |d|", worst, "strict share", strict / ref.size)
    assert bad == 0
    assert strict / ref.size > 0.3


def test_quickvina2_large_ligand_parity(ctx, large_box):
    import vdock
    w = large_box
    ctx.grid_upload(w["box"], w["maps"])
    b, e, _ = w["box"].as_tuple()
    chains, depth, bfgs = 128, 300, 16
    dev, ref, names = [], [], []
    for heavy, tors, lseed in ((44, 12, 901), (40, 14, 902)):
        text = vdock.synth_ligand(heavy, tors, lseed)
        lid = ctx.add_ligand(text)
        lig = orc.Ligand(text)
        assert lig.n_heavy >= 40 and lig.T >= 12
        runs = {}
        for policy, seeds in ((1, 16), (0, 4)):
            rd, ro, ev_d, ev_o = [], [], 0, 0
            for seed in range(seeds):
                say(heavy, tors, "policy", policy, "seed", seed)
                eg, _, hg, st = ctx.dock(lid, vdock.default_params(chains=chains, search_depth=depth, bfgs_steps=bfgs,
                                                                   seed=seed, policy=policy))
                eo, _, ho, ev = orc.mc_run(w["shared"], lig, orc.mc_params(chains, depth, bfgs, b, e, seed=seed,
                                                                           stream=lid, threads=16, policy=policy))
                rd.append((float(eg.min()), hg[int(np.argmin(eg))].astype(np.float64)))
                ro.append((float(eo.min()), ho[int(np.argmin(eo))]))
                ev_d += st.evals
                ev_o += ev
            runs[policy] = (rd, ro, ev_d / seeds, ev_o / seeds)
        r_d = runs[1][2] / runs[0][2]
        r_o = runs[1][3] / runs[0][3]
        say("%d/%d: evaluations of the QuickVina2 policy over the Vina policy: device %.3f oracle %.3f" % (
            lig.n_atoms, lig.T, r_d, r_o))
        assert r_o < 0.9                     # the histories fill and the gate skips searches
        assert abs(r_d - r_o) <= 0.05
        dev.append(runs[1][0])
        ref.append(runs[1][1])
        names.append("%d/%d" % (lig.n_atoms, lig.T))
    assert_parity(multi_ligand_parity(dev, ref, names))


def test_large_box_batch_properties(ctx, large_box):
    import vdock
    from vdock import workloads as W
    w = large_box
    ctx.grid_upload(w["box"], w["maps"])
    specs = W.large_box_specs()
    texts = W.texts(vdock, specs)
    say("large box: 4096 ligands")
    ids = [ctx.add_ligand(t) for t in texts]
    p = vdock.default_params(chains=128, seed=11, policy=vdock.VD_MC_QVINA2)
    st = ctx.dock_batch(ids, p)
    say("large box batch: kernel ms", st.kernel_ms, "evals", st.evals)
    best = ctx.batch_best()
    assert np.isfinite(best).all()
    assert {r.lanes_per_chain for r in ctx.last_launches()} == {16, 64}
    tors = [ctx.ligand_info(i).n_torsions for i in ids]
    assert max(tors) >= 20
    worst = 0.0
    # a sample spanning the sizes: every 128th ligand and the one with the most torsions
    sample = sorted(set(range(0, len(ids), 128)) | {int(np.argmax(tors))})
    for k in sample:
        e, conf, heavy = ctx.batch_results(k)
        assert len(e) == 128
        worst = max(worst, check_best_chains(ctx, w["shared"], orc.Ligand(texts[k]), ids[k], e, conf, heavy, w["box"], k_best=4))
    say("large box: ligands checked", len(sample), "worst |e - e_oracle|", worst)
