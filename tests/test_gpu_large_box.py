"""The large-box screening workload as a parity case: 20,000-atom synthetic receptor (90 A cube,
25 A cavity), 70 A box at 0.375 A (188^3 points, 17 FP32 maps = 452 MB), the QuickVina2 policy.

  - every point of the 17 device maps against the FP64 oracle's cache::populate:
    |d| <= 1e-4 max(1, |ref|) + slack (the oracle's bound on what an FP32 rounding of r^2 within
    0.003 LUT bins of a bin edge or the cutoff can change);
  - the 4,096-ligand batch (10-80 heavy atoms, 0-24 torsions drawn, 128 chains each, QuickVina2
    policy, one persistent launch per lanes-per-chain class): finite results for every ligand,
    and on a sample of ligands the best chains' energies are the oracle's evaluation of the
    returned conformations on the same maps and their heavy-atom coordinates those of the
    conformation.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import check_best_chains, maps_within, say

pytestmark = pytest.mark.gpu


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
        worst = max(worst, check_best_chains(ctx, w["shared"], orc.Ligand(texts[k]), ids[k], e, conf, heavy, k_best=4))
    say("large box: ligands checked", len(sample), "worst |e - e_oracle|", worst)
