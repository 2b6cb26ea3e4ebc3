"""The dense-BFGS local search (the reference's --rilc_bfgs 0: kernel2.cu:427-563 with
bfgs_optimized, quasi_newton.cuh:2391-2499; CPU form bfgs.h:30-140) on the device.

  - docking parity over 16 seeds against the oracle's Monte Carlo with dense BFGS, on the
    single-ligand scene (8,000 atoms, 81^3 maps), two ligands (20 heavy atoms / 4 torsions,
    32 / 8), 256 chains, the dense-BFGS search-depth and step heuristics (V11/main/main.cpp:354-391):
    helpers.multi_ligand_parity;
  - the batched persistent launch and the plain per-ligand launches run the same chains: with the
    lanes per chain fixed, every chain's energy, conformation and coordinates are equal bit for
    bit (a chain's trajectory depends only on its ligand, chain index and seed).
"""
import numpy as np
import pytest

import oracle as orc
from helpers import assert_parity, multi_ligand_parity, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(module_ctx):
    """This module's own context (conftest.module_ctx): its random streams do not depend on the
    tests that ran before it."""
    return module_ctx


@pytest.fixture(scope="module")
def scene(ctx):
    import vdock
    from vdock import workloads as W
    say("dense BFGS: single-ligand scene")
    xyz, xs, box = W.single_ligand_scene(vdock, ctx)
    maps = ctx.grid_download()
    shared = orc.Grid(orc.Lut(), *box.as_tuple())
    shared.set_maps(maps.astype(np.float64))
    return dict(box=box, maps=maps, shared=shared)


def test_dense_bfgs_docking_parity(ctx, scene):
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    b, e, _ = scene["box"].as_tuple()
    dev, ref, names = [], [], []
    for heavy, tors, lseed in ((20, 4, 801), (32, 8, 802)):
        text = vdock.synth_ligand(heavy, tors, lseed)
        lid = ctx.add_ligand(text)
        lig = orc.Ligand(text)
        depth, bfgs = vdock.heuristics(lig.n_atoms, lig.T, rilc=False)
        rd, ro = [], []
        for seed in range(16):
            say("dense", heavy, tors, "seed", seed)
            eg, _, hg, st = ctx.dock(lid, vdock.default_params(chains=256, seed=seed, local_search=vdock.VD_LS_BFGS))
            assert [r.mode for r in ctx.last_launches()] == [0]
            eo, _, ho, _ = orc.mc_run(scene["shared"], lig, orc.mc_params(256, depth, bfgs, b, e, seed=seed, stream=lid,
                                                                          threads=16, mode=0))
            rd.append((float(eg.min()), hg[int(np.argmin(eg))].astype(np.float64)))
            ro.append((float(eo.min()), ho[int(np.argmin(eo))]))
        dev.append(rd)
        ref.append(ro)
        names.append("%d/%d" % (lig.n_atoms, lig.T))
    assert_parity(multi_ligand_parity(dev, ref, names))


def test_dense_bfgs_batched_equals_plain(ctx, scene):
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    ids = [ctx.add_ligand(vdock.synth_ligand(12 + 5 * k, (2, 5, 9, 3)[k], 810 + k)) for k in range(4)]
    p = vdock.default_params(chains=96, search_depth=4, bfgs_steps=10, seed=3, local_search=vdock.VD_LS_BFGS,
                             chains_per_wave=2)
    out = {}
    for plain in (False, True):
        say("dense batched" if not plain else "dense plain")
        ctx.dock_batch(ids, p, plain=plain)
        assert {r.mode for r in ctx.last_launches()} == {0}
        out[plain] = [ctx.batch_results(k) for k in range(len(ids))]
    for k in range(len(ids)):
        for x, y in zip(out[False][k], out[True][k]):
            assert np.isfinite(x).all()
            assert np.array_equal(x, y), k
