"""The batched screening workload (the headline's path) as a parity case: the single-ligand scene's
receptor and maps (8,000 atoms, 30 A box, 81^3 points), synthetic ligands of 10-60 heavy atoms and
0..min(16, atoms/3) torsions, 256 chains each, RILC-BFGS, one persistent launch per lanes-per-chain
class with the two classes running at the same time.

  - twelve of the screening ligands (eight of the 16-lane class, four of the 64-lane class),
    batched together as the headline step batches them, docked over 16 seeds against the oracle's
    Monte Carlo with the same schedule, heuristics and random stream: statistical parity as
    helpers.multi_ligand_parity states it (per-ligand Mann-Whitney on the runs' best scores,
    pooled mean best score within 0.1 kcal/mol, mean best-of-16 no more than 0.1 above the
    oracle's, top-pose successes not significantly fewer);
  - the whole 1,024-ligand batch: finite results for every ligand, and on a sample of ligands the
    best chains' energies are the oracle's evaluation of the returned conformations on the same
    maps, their heavy-atom coordinates those of the conformation.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import assert_parity, check_best_chains, multi_ligand_parity, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(module_ctx):
    """This module's own context (conftest.module_ctx): its random streams do not depend on the
    tests that ran before it."""
    return module_ctx


@pytest.fixture(scope="module")
def screening(ctx):
    import vdock
    from vdock import workloads as W
    say("screening scene: 8000-atom receptor, 81^3 maps")
    xyz, xs, box = W.single_ligand_scene(vdock, ctx)
    maps = ctx.grid_download()
    shared = orc.Grid(orc.Lut(), *box.as_tuple())
    shared.set_maps(maps.astype(np.float64))
    return dict(xyz=xyz, xs=xs, box=box, maps=maps, shared=shared)


# screening ligands spanning both lanes-per-chain classes (atoms, torsions): 22/5, 29/2, 60/6,
# 33/7, 37/6, 30/3, 56/4, 42/4 in the 16-lane class; 54/10, 46/11, 46/13, 35/11 in the 64-lane one
PARITY_SUBSET = [0, 2, 8, 10, 13, 19, 22, 33, 11, 17, 23, 37]


def test_batched_screening_parity_over_16_seeds(ctx, screening):
    import vdock
    from vdock import workloads as W
    w = screening
    ctx.grid_upload(w["box"], w["maps"])
    specs = W.screening_specs()
    texts = [vdock.synth_ligand(*specs[i]) for i in PARITY_SUBSET]
    ids = [ctx.add_ligand(t) for t in texts]
    ligs = [orc.Ligand(t) for t in texts]
    b, e, _ = w["box"].as_tuple()
    dev = [[] for _ in ids]
    ref = [[] for _ in ids]
    for seed in range(16):
        say("screening subset, seed", seed)
        ctx.dock_batch(ids, vdock.default_params(chains=256, seed=seed))
        assert sorted({r.lanes_per_chain for r in ctx.last_launches()}) == [16, 64]
        for k, (lid, lig) in enumerate(zip(ids, ligs)):
            eg, _, hg = ctx.batch_results(k)
            dev[k].append((float(eg.min()), hg[int(np.argmin(eg))].astype(np.float64)))
            d, bf = vdock.heuristics(lig.n_atoms, lig.T)
            eo, _, ho, _ = orc.mc_run(w["shared"], lig, orc.mc_params(256, d, bf, b, e, seed=seed, stream=lid, threads=16))
            ref[k].append((float(eo.min()), ho[int(np.argmin(eo))]))
    names = ["%d/%d" % (lig.n_atoms, lig.T) for lig in ligs]
    assert_parity(multi_ligand_parity(dev, ref, names))


def test_screening_batch_properties(ctx, screening):
    import vdock
    from vdock import workloads as W
    w = screening
    ctx.grid_upload(w["box"], w["maps"])
    specs = W.screening_specs()
    texts = W.texts(vdock, specs)
    say("screening: 1024 ligands")
    ids = [ctx.add_ligand(t) for t in texts]
    p = vdock.default_params(chains=256, seed=7)
    st = ctx.dock_batch(ids, p)
    launches = ctx.last_launches()
    say("screening batch: kernel ms", st.kernel_ms, "evals", st.evals, [r.kernel for r in launches])
    assert sorted(r.lanes_per_chain for r in launches) == [16, 64]
    assert launches[1].start_ms < launches[0].start_ms + launches[0].kernel_ms   # the classes overlap
    best = ctx.batch_best()
    assert np.isfinite(best).all()
    worst = 0.0
    sample = list(range(0, len(ids), 64))
    for k in sample:
        e, conf, heavy = ctx.batch_results(k)
        assert len(e) == 256
        worst = max(worst, check_best_chains(ctx, w["shared"], orc.Ligand(texts[k]), ids[k], e, conf, heavy, w["box"]))
    say("screening: ligands checked", len(sample), "worst |e - e_oracle|", worst)
