"""Edge cases of the docking entry points on the device, each against the oracle or against the
path's own definition:

  - an empty batch: no launch, no evaluations, nothing to fetch;
  - a chain count that leaves the last chain group of a wavefront partly empty (37 chains at four
    chains per wavefront): the plain and the batched launch return the same 37 chains, and the
    best chains' energies are the oracle's evaluation of their conformations;
  - a rigid ligand (no torsion) and a ligand with many torsions, batched together with a typical
    one: finite results, best chains confirmed by the oracle;
  - the same ligand twice in one batch: each slot's chains draw from the ligand's own random stream
    (keyed by the seed and the ligand id), so the two slots return the same chains bit for bit;
  - a ligand wider than the largest group (120 atoms, 24 torsions) docked through the plain path:
    finite results, best chains confirmed by the oracle.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import check_best_chains, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene(ctx):
    import vdock
    xyz, xs = vdock.synth_receptor(2000, 36.0, 8.0, 42)
    box = vdock.box_from_center([0, 0, 0], [20.0] * 3, 0.375)
    ctx.set_receptor(xyz, xs)
    ctx.grid_populate(box)
    maps = ctx.grid_download()
    g = orc.Grid(orc.Lut(), *box.as_tuple())
    g.set_maps(maps.astype(np.float64))
    return dict(box=box, maps=maps, grid=g)


def test_empty_batch(ctx, scene):
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    st = ctx.dock_batch([], vdock.default_params(chains=64))
    assert st.evals == 0 and st.blocks == 0
    assert ctx.last_launches() == []
    assert len(ctx.batch_best()) == 0


def test_partial_chain_group(ctx, scene):
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    text = vdock.synth_ligand(18, 3, 21)
    lid = ctx.add_ligand(text)
    lig = orc.Ligand(text)
    p = vdock.default_params(chains=37, search_depth=4, bfgs_steps=10, seed=2, chains_per_wave=4)
    e, conf, heavy, st = ctx.dock(lid, p)
    assert e.shape == (37,) and st.chains_per_wave == 4
    ctx.dock_batch([lid], p)
    eb, cb, hb = ctx.batch_results(0)
    assert np.array_equal(e, eb) and np.array_equal(conf, cb) and np.array_equal(heavy, hb)
    worst = check_best_chains(ctx, scene["grid"], lig, lid, e, conf, heavy, scene["box"])
    say("37 chains: worst |e - e_oracle|", worst)


def test_rigid_and_many_torsions_in_one_batch(ctx, scene):
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    texts = [vdock.synth_ligand(16, 0, 31), vdock.synth_ligand(45, 15, 32), vdock.synth_ligand(24, 5, 33)]
    ids = [ctx.add_ligand(t) for t in texts]
    assert [ctx.ligand_info(i).n_torsions for i in ids][0] == 0
    assert ctx.ligand_info(ids[1]).n_torsions >= 12
    st = ctx.dock_batch(ids, vdock.default_params(chains=128, seed=4))
    assert st.evals > 0
    for k, (lid, t) in enumerate(zip(ids, texts)):
        e, conf, heavy = ctx.batch_results(k)
        assert len(e) == 128
        worst = check_best_chains(ctx, scene["grid"], orc.Ligand(t), lid, e, conf, heavy, scene["box"])
        say("ligand", k, ctx.ligand_info(lid).n_torsions, "torsions: worst |e - e_oracle|", worst)


def test_same_ligand_twice_in_a_batch(ctx, scene):
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    lid = ctx.add_ligand(vdock.synth_ligand(20, 4, 41))
    other = ctx.add_ligand(vdock.synth_ligand(40, 12, 42))
    ctx.dock_batch([lid, other, lid], vdock.default_params(chains=96, search_depth=3, bfgs_steps=8, seed=6))
    a, b = ctx.batch_results(0), ctx.batch_results(2)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_wide_ligand_plain_launch(ctx, scene):
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    text = vdock.synth_ligand(120, 24, 51)
    lid = ctx.add_ligand(text)
    inf = ctx.ligand_info(lid)
    assert inf.n_atoms > 64 and inf.n_torsions >= 20
    lig = orc.Ligand(text)
    e, conf, heavy, st = ctx.dock(lid, vdock.default_params(chains=128, search_depth=2, bfgs_steps=6, seed=8))
    assert st.evals > 0
    worst = check_best_chains(ctx, scene["grid"], lig, lid, e, conf, heavy, scene["box"])
    say("120-atom ligand: worst |e - e_oracle|", worst)
