"""The batched screening workload (the headline's path) as a parity case: the single-ligand scene's
receptor and maps (8,000 atoms, 30 A box, 81^3 points), synthetic ligands of 10-60 heavy atoms and
0..min(16, atoms/3) torsions, 256 chains each, RILC-BFGS, one persistent launch per lanes-per-chain
class with the two classes running at the same time.

  - the whole 1,024-ligand batch: finite results for every ligand, and on a sample of ligands the
    best chains' energies are the oracle's evaluation of the returned conformations on the same
    maps, their heavy-atom coordinates those of the conformation.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import check_best_chains, say

pytestmark = pytest.mark.gpu


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
        worst = max(worst, check_best_chains(ctx, w["shared"], orc.Ligand(texts[k]), ids[k], e, conf, heavy))
    say("screening: ligands checked", len(sample), "worst |e - e_oracle|", worst)
