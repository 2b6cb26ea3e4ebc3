"""Shared pieces of the GPU parity tests: error bounds derived from FP32 rounding, pose
distances, and the statistics that compare the device's docking runs with the oracle's."""
import numpy as np

import oracle as orc


def say(*a):
    print(*a, flush=True)


def rmsd(a, b):
    return float(np.sqrt(((np.asarray(a) - np.asarray(b)) ** 2).sum(axis=-1).mean()))


def out_of_box_bound(lig, box, conf, slope=1e6, ulps=8):
    """The out-of-box penalty is slope x distance, not curled: FP32 coordinates (a few roundings
    of the kinematics, `ulps` units in the last place) move it by slope x that much per axis on
    which a scored atom lies outside the box or within that rounding of its wall (the penalty is
    continuous at the wall, so a coordinate error d changes it by at most slope x d either way)."""
    b, e, _ = (np.asarray(v, dtype=np.float64) for v in box.as_tuple())
    X = lig.set_conf(np.asarray(conf, dtype=np.float64))[lig.atoms()["xs"] < orc.N_XS]
    delta = ulps * np.spacing(np.abs(X).astype(np.float32)).astype(np.float64)
    near = (X < b + delta) | (X > e - delta)
    return float((slope * delta * near).sum())


def map_error_bounds(lig, box, tol, conf):
    """What map errors of at most `tol` (per map point) can do to a pose's energy and gradient:
    the energy moves by at most the trilinear interpolation of tol at each scored atom (curl only
    shrinks differences); an atom's force by at most sqrt(3)/granularity times the largest tol of
    its cell's corners; a gradient component -- a sum of forces, or of torques about an atom of
    the ligand -- by at most the sum of those times max(1, ligand diameter)."""
    b, e, n = (np.asarray(v, dtype=np.float64) for v in box.as_tuple())
    factor = n / (e - b)
    at = lig.atoms()
    X = lig.set_conf(np.asarray(conf, dtype=np.float64))
    diam = max(1.0, float(np.sqrt(((X[:, None] - X[None]) ** 2).sum(-1)).max()))
    eb = fb = 0.0
    for i in np.nonzero(at["xs"] < orc.N_XS)[0]:
        s = np.clip((X[i] - b) * factor, 0, n)
        a = np.minimum(np.floor(s).astype(int), n.astype(int) - 1)
        f = s - a
        c = tol[at["xs"][i], a[2]:a[2] + 2, a[1]:a[1] + 2, a[0]:a[0] + 2]
        wz, wy, wx = (np.array([1 - f[k], f[k]]) for k in (2, 1, 0))
        eb += float((c * wz[:, None, None] * wy[None, :, None] * wx[None, None, :]).sum())
        fb += np.sqrt(3) * factor.max() * float(c.max())
    return eb, fb * diam


_CUTOFF_JUMP = None


def intra_cutoff_bound(lig, conf, ulps=8):
    """The pair table is not zero at the 8 A cutoff (the gauss2 term of two carbons is about
    -0.0036 kcal/mol there) and a pair at or beyond it contributes nothing: an intramolecular pair
    whose r^2 lies within FP32 rounding of 64 (coordinates rounded by `ulps` units in the last
    place, r^2 = 64 +- 2 r (2 d)) may be counted on one side and not on the other, which moves the
    energy by the table's value there."""
    global _CUTOFF_JUMP
    if _CUTOFF_JUMP is None:
        lut = orc.Lut()
        _CUTOFF_JUMP = np.array([np.abs(lut.tables(tp)[0][2046:2051]).max() for tp in range(orc.N_PAIRS)])
    X = lig.set_conf(np.asarray(conf, dtype=np.float64))
    pairs = lig.pairs()
    if len(pairs) == 0:
        return 0.0
    d = ulps * float(np.spacing(np.float32(np.abs(X).max())))
    r2 = ((X[pairs[:, 1]] - X[pairs[:, 2]]) ** 2).sum(-1)
    near = np.abs(r2 - 64.0) <= 2 * 8.0 * 2 * d
    return float(_CUTOFF_JUMP[pairs[near, 0]].sum())


def eval_bound(e_ref, g_ref):
    """The evaluation's own bound on shared maps (the parity target: 1e-4 kcal/mol absolute / 1e-3
    relative on energies; gradients per component 1e-3 of the largest component plus 1e-4)."""
    return max(1e-4, 1e-3 * abs(e_ref)), 1e-3 * max(1.0, float(np.abs(g_ref).max())) + 1e-4


def maps_within(dev, ref, slack):
    """Device maps against the oracle's, map by map (large boxes): |d| <= 1e-4 max(1, |ref|) +
    slack, where slack is the oracle's bound on the change an FP32 rounding of r^2 within 0.003
    LUT bins of a bin edge or the cutoff can cause.  Returns (points outside, max |d|, points
    where the bound is strict)."""
    bad, worst, strict = 0, 0.0, 0
    for t in range(dev.shape[0]):
        d = np.abs(dev[t].astype(np.float64) - ref[t])
        tol = 1e-4 * np.maximum(1.0, np.abs(ref[t])) + slack[t]
        bad += int((d > tol).sum())
        worst = max(worst, float(d.max()))
        strict += int((slack[t] == 0).sum())
    return bad, worst, strict


def check_best_chains(ctx, grid, lig, lid, e, conf, heavy, box, k_best=8, rel=False):
    """Property checks of one ligand's docking output: every value finite; the energies of the k
    best chains are the oracle's evaluation (search caps: v = 10, slope 1e6) of the returned
    conformations on the same maps, within max(1e-3, 1e-3 |e|) plus what FP32 rounding can do to
    the two discontinuous parts of the energy: the out-of-box penalty (poses the search presses
    against a wall of the box) and intramolecular pairs at the cutoff; the heavy-atom coordinates
    are those of the returned conformation.  Returns the worst |e - e_oracle|, or with rel the worst
    |e - e_oracle| / max(1, |e_oracle|)."""
    assert np.isfinite(e).all() and np.isfinite(conf).all() and np.isfinite(heavy).all()
    worst = 0.0
    for i in np.argsort(e)[:k_best]:
        c = conf[i].astype(np.float64)
        e_ref = orc.eval_deriv(grid, lig, c)[0]
        base = max(1e-3, 1e-3 * abs(e_ref))
        ob, ic = out_of_box_bound(lig, box, c), intra_cutoff_bound(lig, c)
        assert abs(e[i] - e_ref) <= base + ob + ic, (lid, i, e[i], e_ref)
        if abs(e[i] - e_ref) > base:
            b0, e0, _ = (np.asarray(v, dtype=np.float64) for v in box.as_tuple())
            X = lig.set_conf(c)[lig.atoms()["xs"] < orc.N_XS]
            out = np.maximum(b0 - X, 0) + np.maximum(X - e0, 0)
            say(f"chain {i}: |e - e_oracle| {abs(e[i] - e_ref):.4g} over the base {base:.3g}, inside the out-of-box "
                f"bound {ob:.4g} + cutoff bound {ic:.3g}; oracle penalty {1e6 * out.sum():.4g} from "
                f"{int((out > 0).any(1).sum())} atoms outside the box")
        worst = max(worst, abs(e[i] - e_ref) / (max(1.0, abs(e_ref)) if rel else 1.0))
        x = lig.set_conf(c)[lig.atoms()["el"] != 0]
        assert np.abs(x - heavy[i]).max() < 2e-3, (lid, i)
    return worst


def no_worse_success(k_dev, k_orc, n, alpha=0.01):
    """One-sided Fisher's exact test: is the device's count of successful runs (k_dev of n)
    significantly below the oracle's (k_orc of n)?  True when it is not, at level alpha."""
    from scipy.stats import fisher_exact
    p = fisher_exact([[k_dev, n - k_dev], [k_orc, n - k_orc]], alternative="less").pvalue
    say("success", k_dev, "of", n, "vs oracle", k_orc, "of", n, "one-sided p", p)
    return p >= alpha


def multi_ligand_parity(dev, orc_, names, mw_alpha=1e-3, mean_bound=0.1, best_bound=0.1, boot=2000):
    """Statistical docking parity over several ligands.  dev / orc_: per ligand, a list over seeds
    of (best energy, top pose heavy-atom coordinates) of one run.  Two implementations of the same
    search are two samples of the same distribution of run outcomes, so:
      - per ligand, the runs' best scores pass a two-sided Mann-Whitney test at level mw_alpha (a
        shift either way is caught; 12 ligands at 1e-3 keep the family-wise false alarm near 1 %);
      - the mean per-run best score, averaged over ligands, differs by at most mean_bound kcal/mol
        (the 0.1 kcal/mol target) beyond three standard errors of its measurement (bootstrap over
        each side's runs): a violation is declared only where the seeds resolve it;
      - the best score over all seeds (the docking target's "best score"), averaged over ligands, is
        at most best_bound above the oracle's, beyond three bootstrap standard errors; per ligand
        it scatters as much between two seed sets of the oracle itself as between the two
        implementations;
      - top-pose RMSD: a run succeeds when its top pose lies within 2 A of the best pose found by
        either side; the device's successes, pooled over ligands, are not significantly fewer
        than the oracle's (Fisher, one-sided, 1 %)."""
    from scipy.stats import mannwhitneyu
    rng = np.random.default_rng(0)
    mean_diffs, best_diffs, k_dev, k_orc, n = [], [], 0, 0, 0
    boot_mean, boot_best = np.zeros(boot), np.zeros(boot)
    ok = True
    for name, d, o in zip(names, dev, orc_):
        ed, eo = np.array([x[0] for x in d]), np.array([x[0] for x in o])
        p = mannwhitneyu(ed, eo).pvalue
        ref = (d + o)[int(np.argmin(np.concatenate([ed, eo])))][1]
        sd = sum(rmsd(x[1], ref) <= 2.0 for x in d)
        so = sum(rmsd(x[1], ref) <= 2.0 for x in o)
        k_dev, k_orc, n = k_dev + sd, k_orc + so, n + len(d)
        mean_diffs.append(ed.mean() - eo.mean())
        best_diffs.append(ed.min() - eo.min())
        bd = ed[rng.integers(0, len(ed), (boot, len(ed)))]
        bo = eo[rng.integers(0, len(eo), (boot, len(eo)))]
        boot_mean += bd.mean(1) - bo.mean(1)
        boot_best += bd.min(1) - bo.min(1)
        say(name, "best %.3f %.3f mean %.3f %.3f Mann-Whitney p %.3f success %d %d of %d" % (
            ed.min(), eo.min(), ed.mean(), eo.mean(), p, sd, so, len(d)))
        ok = ok and p >= mw_alpha
    L = len(names)
    mean_diff, best_diff = float(np.mean(mean_diffs)), float(np.mean(best_diffs))
    se_mean, se_best = float((boot_mean / L).std()), float((boot_best / L).std())
    say("mean per-run best difference %.4f (se %.4f), mean best-of-seeds difference %.4f (se %.4f)" % (
        mean_diff, se_mean, best_diff, se_best))
    return dict(mw_ok=ok, mean_diff=mean_diff, best_diff=best_diff, se_mean=se_mean, se_best=se_best,
                success_ok=no_worse_success(k_dev, k_orc, n), mean_bound=mean_bound, best_bound=best_bound)


def assert_parity(r):
    assert r["mw_ok"]
    assert abs(r["mean_diff"]) <= r["mean_bound"] + 3 * r["se_mean"]
    assert r["best_diff"] <= r["best_bound"] + 3 * r["se_best"]
    assert r["success_ok"]


def compare_chain_distributions(name, G, O, gb, ob, gev, oev, check=True):
    """Two docking runs compared chain by chain: G, O = per-run arrays of every chain's final
    energy, gb, ob = per-run best energies, gev, oev = evaluation counts.  Criteria: two-sample
    Kolmogorov-Smirnov p >= 1e-3 on the pooled chains, the 1/5/25/50 % quantiles within 0.05
    kcal/mol, the mean per-run best within 3.5 standard errors of the difference, the evaluation
    counts within 1 %.  Returns the figures; asserts them unless check is False."""
    from scipy.stats import ks_2samp
    G, O = np.concatenate(G), np.concatenate(O)
    gb, ob = np.asarray(gb), np.asarray(ob)
    qs = [0.01, 0.05, 0.25, 0.5]
    qg, qo = np.quantile(G, qs), np.quantile(O, qs)
    ks = ks_2samp(G, O).pvalue
    se = np.sqrt(gb.var(ddof=1) / len(gb) + ob.var(ddof=1) / len(ob))
    r = dict(ks=ks, dq=float(np.abs(qg - qo).max()), dbest=float(abs(gb.mean() - ob.mean())), se=float(se),
             devals=abs(gev - oev) / max(oev, 1))
    say(name, "chains", len(G), "KS p", ks, "quantiles", np.round(qg, 4), np.round(qo, 4),
        "mean best", gb.mean(), ob.mean(), "se", se, "evals", gev, oev)
    r["ok"] = bool(ks >= 1e-3 and r["dq"] <= 0.05 and r["dbest"] <= 3.5 * se and r["devals"] <= 0.01)
    if check:
        assert ks >= 1e-3, (name, ks)
        assert r["dq"] <= 0.05, (name, qg, qo)
        assert r["dbest"] <= 3.5 * se, (name, gb.mean(), ob.mean(), se)
        assert r["devals"] <= 0.01, (name, gev, oev)
    return r
