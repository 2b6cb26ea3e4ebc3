"""Edge cases of the docking entry points on the device, each against the oracle or against the
path's own definition:

  - an empty batch: no launch, no evaluations, nothing to fetch;
  - a chain count that leaves the last chain group of a wavefront partly empty (37 chains at four
    chains per wavefront): the plain and the batched launch return the same 37 chains, and the
    best chains' energies are the oracle's evaluation of their conformations;
  - a rigid ligand (no torsion) and a ligand with many torsions, batched together with a typical
    one, under RILC-BFGS and dense BFGS: finite results, best chains confirmed by the oracle;
  - the same ligand twice in one batch: each slot's chains draw from the ligand's own random stream
    (keyed by the seed and the ligand id), so the two slots return the same chains bit for bit;
  - a ligand wider than the largest group (120 atoms, 24 torsions) docked through the plain path:
    finite results, best chains confirmed by the oracle;
  - the largest ligand the library takes (255 atoms, 8-bit atom fields of the pair words; 57
    torsions, 7 + T numbers on 64 lanes): energies and gradients against the oracle, plain and
    batched docking with best chains confirmed by the oracle (also under the QuickVina2,
    QuickVina-W and dense-BFGS modes, whose per-chain state grows with the torsions), and FP64
    refinement of docked poses against the oracle's;
  - the batched launch against plain per-ligand launches of the same ligands (one per class), Vina
    and QuickVina2 policies, RILC-BFGS and dense BFGS: the same chains bit for bit;
  - one past each limit: 259 atoms and 58 torsions are both refused when the ligand is added,
    with the library's error codes.
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


@pytest.mark.parametrize("local_search", [1, 0])
def test_rigid_and_many_torsions_in_one_batch(ctx, scene, local_search):
    """RILC-BFGS and dense BFGS (whose matrix spans the 6 + T dimensions: 6 for the rigid ligand)."""
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    texts = [vdock.synth_ligand(16, 0, 31), vdock.synth_ligand(45, 15, 32), vdock.synth_ligand(24, 5, 33)]
    ids = [ctx.add_ligand(t) for t in texts]
    assert [ctx.ligand_info(i).n_torsions for i in ids][0] == 0
    assert ctx.ligand_info(ids[1]).n_torsions >= 12
    st = ctx.dock_batch(ids, vdock.default_params(chains=128, seed=4, local_search=local_search))
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


@pytest.mark.parametrize("policy,local_search", [(0, 1), (1, 1), (0, 0)])
def test_batched_equals_plain_bit_for_bit(ctx, scene, policy, local_search):
    """A plain per-ligand launch given the lanes per chain of its ligand's class (a plain launch of
    few chains otherwise takes one chain per wavefront, for parallelism) runs the chains as the
    batched launch does: the same lanes per chain and register budget, and each chain drawing from
    the ligand's own random stream.  So the two return the same chains bit for bit -- for the Vina
    policy and for QuickVina2, whose histories are per chain (depth 64: the gate skips searches),
    with RILC-BFGS, and for the Vina policy with dense BFGS.
    One ligand in each class: 20 atoms / 2 torsions (16 lanes), 40 atoms / 12 torsions (64 lanes)."""
    import vdock
    ctx.grid_upload(scene["box"], scene["maps"])
    ids = [ctx.add_ligand(vdock.synth_ligand(20, 2, 81)), ctx.add_ligand(vdock.synth_ligand(40, 12, 82))]
    kw = dict(chains=64, search_depth=64 if local_search else 8, bfgs_steps=8, seed=10, policy=policy,
              local_search=local_search)
    p = vdock.default_params(**kw)
    plain = [ctx.dock(i, vdock.default_params(chains_per_wave=cpw, **kw)) for i, cpw in zip(ids, (4, 1))]
    assert [r[3].chains_per_wave for r in plain] == [4, 1]
    st = ctx.dock_batch(ids, p)
    assert sorted(r.lanes_per_chain for r in ctx.last_launches()) == [16, 64]
    assert st.evals == sum(r[3].evals for r in plain)
    for k, (e, conf, heavy, _) in enumerate(plain):
        eb, cb, hb = ctx.batch_results(k)
        assert np.array_equal(e, eb) and np.array_equal(conf, cb) and np.array_equal(heavy, hb), k


@pytest.fixture(scope="module")
def big_scene(ctx):
    """A receptor shell with a cavity of 14 A radius and a 30 A box, which the largest ligand
    (radius 11-12 A about its centre) fits inside."""
    import vdock
    xyz, xs = vdock.synth_receptor(3000, 50.0, 14.0, 43)
    box = vdock.box_from_center([0, 0, 0], [30.0] * 3, 0.375)
    ctx.set_receptor(xyz, xs)
    ctx.grid_populate(box)
    maps = ctx.grid_download()
    g = orc.Grid(orc.Lut(), *box.as_tuple())
    g.set_maps(maps.astype(np.float64))
    g.set_receptor(xyz.astype(np.float64), xs)
    text = vdock.synth_ligand(236, 57, 6)
    return dict(xyz=xyz, xs=xs, box=box, maps=maps, grid=g, text=text, lig=orc.Ligand(text))


def largest_ligand(ctx, s):
    ctx.set_receptor(s["xyz"], s["xs"])
    ctx.grid_upload(s["box"], s["maps"])
    lid = ctx.add_ligand(s["text"])
    inf = ctx.ligand_info(lid)
    assert (inf.n_atoms, inf.n_torsions) == (255, 57)
    return lid


def test_largest_ligand_eval(ctx, big_scene):
    from helpers import intra_cutoff_bound, out_of_box_bound
    s = big_scene
    lid, lig = largest_ligand(ctx, s), s["lig"]
    say("largest ligand: pairs", lig.n_pairs)
    rng = np.random.default_rng(5)
    n = 32
    conf = np.zeros((n, 7 + lig.T))
    conf[:, :3] = rng.uniform(-2.0, 2.0, (n, 3))
    q = rng.normal(size=(n, 4))
    conf[:, 3:7] = q / np.linalg.norm(q, axis=1, keepdims=True)
    conf[:, 7:] = rng.uniform(-np.pi, np.pi, (n, lig.T))
    conf = conf.astype(np.float32)
    e_gpu, g_gpu = ctx.eval(lid, conf)   # 7 + 57 numbers: one chain per wavefront
    worst_e = worst_g = 0.0
    for i in range(n):
        c = conf[i].astype(np.float64)
        e_ref, g_ref, _, _ = orc.eval_deriv(s["grid"], lig, c)
        tol_e = max(1e-3, 1e-3 * abs(e_ref)) + intra_cutoff_bound(lig, c) + out_of_box_bound(lig, s["box"], c)
        assert abs(e_gpu[i] - e_ref) <= tol_e, (i, e_gpu[i], e_ref)
        tol_g = 1e-3 * max(1.0, np.abs(g_ref).max()) + 1e-4
        assert np.all(np.abs(g_gpu[i] - g_ref) <= tol_g), (i, np.abs(g_gpu[i] - g_ref).max())
        worst_e = max(worst_e, abs(e_gpu[i] - e_ref) / max(1.0, abs(e_ref)))
        worst_g = max(worst_g, np.abs(g_gpu[i] - g_ref).max() / max(1.0, np.abs(g_ref).max()))
    say("largest ligand eval: worst rel e", worst_e, "worst rel g", worst_g)


def test_largest_ligand_docks_plain_and_batched(ctx, big_scene):
    import vdock
    s = big_scene
    lid = largest_ligand(ctx, s)
    text_small = vdock.synth_ligand(20, 4, 61)
    small = ctx.add_ligand(text_small)
    p = vdock.default_params(chains=64, search_depth=2, bfgs_steps=6, seed=12)
    e, conf, heavy, st = ctx.dock(lid, p)
    assert st.evals > 64
    worst = check_best_chains(ctx, s["grid"], s["lig"], lid, e, conf, heavy, s["box"], rel=True)
    say("largest ligand, plain: best", e.min(), "worst relative |e - e_oracle|", worst)
    ctx.dock_batch([small, lid], p)
    assert sorted(r.kernel.split("<")[1].split(",")[0] for r in ctx.last_launches()) == ["16", "64"]
    for k, (i, lig) in enumerate([(small, orc.Ligand(text_small)), (lid, s["lig"])]):
        e, conf, heavy = ctx.batch_results(k)
        assert len(e) == 64
        worst = check_best_chains(ctx, s["grid"], lig, i, e, conf, heavy, s["box"], rel=True)
        say("batched ligand", k, "best", e.min(), "worst relative |e - e_oracle|", worst)


@pytest.mark.parametrize("mode", ["quickvina2", "quickvina_w", "dense_bfgs"])
def test_largest_ligand_other_search_modes(ctx, big_scene, mode):
    """The modes whose per-chain state grows with the torsion count, at 57 torsions: the
    QuickVina2 history (a ring of 630 entries of 64 numbers), QuickVina-W's individual ring and the
    ligand's shared cell buffer, the dense BFGS matrix (63 x 63)."""
    import vdock
    s = big_scene
    lid = largest_ligand(ctx, s)
    kw = dict(quickvina2=dict(policy=vdock.VD_MC_QVINA2), quickvina_w=dict(policy=vdock.VD_MC_QVINAW),
              dense_bfgs=dict(local_search=vdock.VD_LS_BFGS))[mode]
    p = vdock.default_params(chains=64, search_depth=3, bfgs_steps=6, seed=14, **kw)
    e, conf, heavy, st = ctx.dock(lid, p)
    assert st.evals > 64
    worst = check_best_chains(ctx, s["grid"], s["lig"], lid, e, conf, heavy, s["box"], rel=True)
    say("largest ligand,", mode, "plain: best", e.min(), "worst relative |e - e_oracle|", worst)
    if mode == "quickvina2":
        ctx.dock_batch([lid], p)
        e, conf, heavy = ctx.batch_results(0)
        worst = check_best_chains(ctx, s["grid"], s["lig"], lid, e, conf, heavy, s["box"], rel=True)
        say("largest ligand,", mode, "batched: best", e.min(), "worst relative |e - e_oracle|", worst)


def test_largest_ligand_refines_like_the_oracle(ctx, big_scene):
    import vdock
    s = big_scene
    lid, lig = largest_ligand(ctx, s), s["lig"]
    en, conf, _, _ = ctx.dock(lid, vdock.default_params(chains=64, search_depth=2, bfgs_steps=6, seed=13))
    starts = conf[np.argsort(en)[:4]].astype(np.float64)
    co, out_xyz, res = ctx.refine(lid, starts)
    de, dr = [], []
    for i, c in enumerate(starts):
        eo, _, xo = orc.refine(s["grid"], lig, c, orc.refine_steps(lig))
        if res[i].inside:
            de.append(abs(res[i].e - eo))
            dr.append(float(np.sqrt(((out_xyz[i].astype(np.float64) - xo) ** 2).sum(-1).mean())))
        else:
            assert eo > 1e300
    say("largest ligand refinement: |de|", np.round(de, 6), "rmsd", np.round(dr, 5))
    assert len(de) >= 3
    assert np.median(de) <= 1e-4 and np.median(dr) <= 1e-3 and max(de) <= 0.1


def test_one_past_each_limit_is_refused(ctx, scene):
    import vdock
    with pytest.raises(vdock.VdockError) as err:
        ctx.add_ligand(vdock.synth_ligand(250, 57, 7))   # 259 atoms
    assert err.value.code == vdock.VD_ERR_PARSE and "limit is 255" in str(err.value)
    with pytest.raises(vdock.VdockError) as err:
        ctx.add_ligand(vdock.synth_ligand(100, 58, 8))   # 105 atoms, 58 torsions
    assert err.value.code == vdock.VD_ERR_CAPACITY and "57 torsions" in str(err.value)
