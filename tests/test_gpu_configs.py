"""The two measured configurations that are parity cases rather than bench lines, at their own
sizes:

  - the CPU-runnable case: 2,000-atom synthetic receptor (36 A cube, 8 A cavity, seed 42), 20 A
    box at 0.375 A (55^3 points), one ligand of 20 heavy atoms and 5 torsions, 512 chains, the
    reference's search_depth / bfgs_steps heuristics -- the device's Monte Carlo against the
    oracle's over 16 seeds, with the criteria of the other docking parity tests;
  - the single-ligand case: 8,000-atom receptor (60 A cube, 10 A cavity), 30 A box (81^3
    points), one ligand of 32 heavy atoms and 8 torsions, 8,192 chains through the plain
    per-ligand path (the drop-in for kernel_monte) -- every result finite, and the oracle
    reproduces the energies of the best chains at their returned conformations on the same maps.
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


def scene(ctx, n_atoms, cube, cavity, seed, box_size):
    import vdock
    xyz, xs = vdock.synth_receptor(n_atoms, cube, cavity, seed)
    ctx.set_receptor(xyz, xs)
    box = vdock.box_from_center([0, 0, 0], [box_size] * 3, 0.375)
    ctx.grid_populate(box)
    maps = ctx.grid_download()
    g = orc.Grid(orc.Lut(), *box.as_tuple())
    g.set_maps(maps.astype(np.float64))          # the oracle runs on the device's maps
    return box, g


def test_cpu_runnable_config_docking_parity(ctx):
    import vdock
    box, g = scene(ctx, 2000, 36.0, 8.0, 42, 20.0)
    assert tuple(box.dims) == (55, 55, 55)
    text = vdock.synth_ligand(20, 5, 1)
    lid = ctx.add_ligand(text)
    inf = ctx.ligand_info(lid)
    lig = orc.Ligand(text)
    depth, bfgs = vdock.heuristics(inf.n_atoms, inf.n_torsions)
    assert (inf.n_heavy, inf.n_torsions, depth, bfgs) == (20, 5, 5, 16)
    b, e, _ = box.as_tuple()
    gb, ob, gp, op = [], [], [], []
    for seed in range(16):
        p = vdock.default_params(chains=512, seed=seed)        # heuristic depth and steps
        eg, _, hg, _ = ctx.dock(lid, p)
        eo, _, ho, _ = orc.mc_run(g, lig, orc.mc_params(512, depth, bfgs, b, e, seed=seed, stream=lid, threads=8))
        gb.append(float(eg.min())); ob.append(float(eo.min()))
        gp.append(hg[int(np.argmin(eg))].astype(np.float64)); op.append(ho[int(np.argmin(eo))])
    assert_parity(multi_ligand_parity([list(zip(gb, gp))], [list(zip(ob, op))], ["config 0: %d/%d" % (lig.n_atoms, lig.T)]))


def test_single_ligand_config_plain_path(ctx):
    import vdock
    box, g = scene(ctx, 8000, 60.0, 10.0, 42, 30.0)
    assert tuple(box.dims) == (81, 81, 81)
    text = vdock.synth_ligand(32, 8, 1)
    lid = ctx.add_ligand(text)
    inf = ctx.ligand_info(lid)
    lig = orc.Ligand(text)
    assert (inf.n_heavy, inf.n_torsions) == (32, 8)
    p = vdock.default_params(chains=8192, seed=11)
    say("config 1: plain launch, 8192 chains")
    e, conf, heavy, st = ctx.dock(lid, p)
    assert e.shape == (8192,) and np.isfinite(e).all() and np.isfinite(conf).all() and np.isfinite(heavy).all()
    assert st.evals > 8192
    for i in np.argsort(e)[:16]:
        e_ref = orc.eval_deriv(g, lig, conf[i].astype(np.float64))[0]
        assert abs(e[i] - e_ref) <= max(1e-3, 1e-3 * abs(e_ref)), (i, e[i], e_ref)
    # the returned heavy-atom coordinates are those of the returned conformation
    i = int(np.argmin(e))
    x = lig.set_conf(conf[i].astype(np.float64))[lig.atoms()["xs"] < orc.N_XS]
    assert np.abs(x - heavy[i]).max() < 1e-3
    say("config 1: best", float(e.min()), "evals", st.evals, "kernel ms", st.kernel_ms)
