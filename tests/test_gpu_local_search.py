"""Deterministic device-vs-oracle checks from identical starting conformations: the local search
alone (vd_local_search) with RILC-BFGS and with dense BFGS, and one Monte Carlo step of vd_dock
from caller-supplied starts (its init_conf, the role of kernel_monte's ric), against the oracle's
local_search and its Monte Carlo from the same starts (orc_mc_run_from).

Scene: the single-ligand workload's receptor and 81^3 maps (the oracle reads the device's maps),
a 24-heavy-atom, 6-torsion synthetic ligand, 512 starting conformations around the box centre
(position within 4 A, uniform orientation, torsions uniform), each exactly representable in FP32.

The device computes in FP32, the oracle in FP64, from the same numbers.  A line search's decisions
(sufficient decrease, curvature) agree unless the quantity compared lies within rounding of its
threshold.  So after one iteration the two must agree start by start: energies within
1e-3 max(1, |e|) and every conformation number within 1e-3 (A, quaternion components, radians)
for at least 95 % of the starts (measured: all of them, with the same evaluation count).  Over a
whole search the two trajectories part -- each iteration's direction update amplifies the
previous difference, and a heuristic step budget stops both before convergence -- so there the
comparison is of distributions: no systematic difference between the device's and the oracle's
end energies (two-sided sign test on the paired differences, p >= 1e-3), the same distribution of
end energies (two-sample Kolmogorov-Smirnov, p >= 1e-3), and the same search effort (evaluation
counts within 5 %).  The same holds for one MC step (mutation, two searches).  Both
lanes-per-chain widths the batched screening uses run (64 lanes, and 16 lanes = four chains per
wavefront).
"""
import numpy as np
import pytest

import oracle as orc
from helpers import say

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
    say("local search: single-ligand scene")
    xyz, xs, box = W.single_ligand_scene(vdock, ctx)
    maps = ctx.grid_download()
    shared = orc.Grid(orc.Lut(), *box.as_tuple())
    shared.set_maps(maps.astype(np.float64))
    text = vdock.synth_ligand(24, 6, 4242)
    lig = orc.Ligand(text)
    rng = np.random.default_rng(17)
    n = 512
    conf = np.zeros((n, 7 + lig.T))
    conf[:, :3] = rng.uniform(-4, 4, (n, 3))
    q = rng.normal(size=(n, 4))
    conf[:, 3:7] = q / np.linalg.norm(q, axis=1, keepdims=True)
    conf[:, 7:] = rng.uniform(-np.pi, np.pi, (n, lig.T))
    conf = conf.astype(np.float32).astype(np.float64)    # the same numbers on both sides
    return dict(box=box, maps=maps, shared=shared, text=text, lig=lig, conf=conf)


def agreement(e_d, c_d, e_o, c_o):
    tol = 1e-3 * np.maximum(1.0, np.abs(e_o))
    de = np.abs(e_d - e_o)
    # quaternions q and -q are the same orientation
    c_d = c_d.copy()
    flip = (c_d[:, 3:7] * c_o[:, 3:7]).sum(1) < 0
    c_d[flip, 3:7] *= -1
    dc = np.abs(c_d - c_o)
    dc[:, 7:] = np.minimum(dc[:, 7:], 2 * np.pi - dc[:, 7:])
    ok = (de <= tol) & (dc.max(1) <= 1e-3)
    return ok, float(np.median(de / np.maximum(1.0, np.abs(e_o))))


def same_distribution(e_d, e_o, ev_d, ev_o):
    """Paired end energies of two searches from the same starts: no systematic difference (sign
    test), the same distribution (Kolmogorov-Smirnov), the same effort (evaluations)."""
    from scipy.stats import binomtest, ks_2samp
    d = e_d - e_o
    nz = d[np.abs(d) > 1e-6 * np.maximum(1.0, np.abs(e_o))]
    p_sign = binomtest(int((nz > 0).sum()), len(nz)).pvalue if len(nz) else 1.0
    p_ks = ks_2samp(e_d, e_o).pvalue
    say("device lower on %d, higher on %d of %d; sign test p %.3g, KS p %.3g, evaluations %d / %d" % (
        int((nz < 0).sum()), int((nz > 0).sum()), len(d), p_sign, p_ks, ev_d, ev_o))
    return p_sign >= 1e-3 and p_ks >= 1e-3 and abs(ev_d - ev_o) <= 0.05 * ev_o


@pytest.mark.parametrize("mode", [1, 0])
@pytest.mark.parametrize("cpw", [1, 4])
def test_local_search_matches_oracle(ctx, scene, mode, cpw):
    import vdock
    s = scene
    ctx.grid_upload(s["box"], s["maps"])
    lid = ctx.add_ligand(s["text"])
    lig = s["lig"]
    e_start = np.array([orc.eval_deriv(s["shared"], lig, c)[0] for c in s["conf"]])
    for steps in (1, 2, 0):
        p = vdock.default_params(local_search=mode, bfgs_steps=steps, chains_per_wave=cpw)
        say("local search mode", mode, "chains per wave", cpw, "steps", steps or "heuristic")
        e_d, c_d, st = ctx.local_search(lid, s["conf"], p)
        assert [r.mode for r in ctx.last_launches()] == [mode] and st.evals >= len(e_d)
        assert np.isfinite(e_d).all() and np.isfinite(c_d).all()
        n_steps = steps or vdock.heuristics(lig.n_atoms, lig.T, rilc=mode == 1)[1]
        e_o, c_o, ev_o = [], [], 0
        for c in s["conf"]:
            e, x, ne = orc.local_search(s["shared"], lig, c, n_steps, mode=mode)
            e_o.append(e)
            c_o.append(x)
            ev_o += ne
        e_o, c_o = np.array(e_o), np.array(c_o)
        ok, med = agreement(e_d.astype(np.float64), c_d.astype(np.float64), e_o, c_o)
        say("agree %.4f  median rel |de| %.2e" % (ok.mean(), med))
        # a search never ends above its start
        assert (e_d <= e_start + 1e-3 * np.maximum(1.0, np.abs(e_start))).all()
        if steps == 1:
            assert ok.mean() >= 0.95 and med <= 1e-4
        elif steps == 0:
            assert same_distribution(e_d.astype(np.float64), e_o, st.evals, ev_o)


@pytest.mark.parametrize("mode", [1, 0])
def test_one_mc_step_from_given_starts_matches_oracle(ctx, scene, mode):
    import vdock
    s = scene
    ctx.grid_upload(s["box"], s["maps"])
    lid = ctx.add_ligand(s["text"])
    lig = s["lig"]
    b, e, _ = s["box"].as_tuple()
    n = len(s["conf"])
    p = vdock.default_params(chains=n, search_depth=1, bfgs_steps=12, seed=9, local_search=mode)
    say("one MC step from given starts, mode", mode)
    e_d, c_d, h_d, st = ctx.dock(lid, p, init_conf=s["conf"])
    e_o, c_o, h_o, ev_o = orc.mc_run(s["shared"], lig, orc.mc_params(n, 1, 12, b, e, seed=9, stream=lid, mode=mode,
                                                                     threads=16), starts=s["conf"])
    ok, med = agreement(e_d.astype(np.float64), c_d.astype(np.float64), e_o, c_o)
    say("agree %.4f  median rel |de| %.2e" % (ok.mean(), med))
    assert np.isfinite(e_d).all()
    assert same_distribution(e_d.astype(np.float64), e_o, st.evals, ev_o)
    # the first step's candidate is the start mutated as the oracle mutates it: the draws and the
    # mutation are the same, so the searches start from the same point and the first iteration agrees
    p1 = vdock.default_params(chains=n, search_depth=1, bfgs_steps=1, seed=9, local_search=mode)
    e1, c1, _, _ = ctx.dock(lid, p1, init_conf=s["conf"])
    o1, oc1, _, _ = orc.mc_run(s["shared"], lig, orc.mc_params(n, 1, 1, b, e, seed=9, stream=lid, mode=mode, threads=16),
                               starts=s["conf"])
    ok1, med1 = agreement(e1.astype(np.float64), c1.astype(np.float64), o1, oc1)
    say("one step, one iteration per search: agree %.4f  median rel |de| %.2e" % (ok1.mean(), med1))
    assert ok1.mean() >= 0.95 and med1 <= 1e-4
