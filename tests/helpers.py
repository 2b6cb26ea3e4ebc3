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
    which a scored atom lies outside the box."""
    b, e, _ = (np.asarray(v, dtype=np.float64) for v in box.as_tuple())
    X = lig.set_conf(np.asarray(conf, dtype=np.float64))[lig.atoms()["xs"] < orc.N_XS]
    out = (X < b) | (X > e)
    return float((slope * ulps * np.spacing(np.abs(X).astype(np.float32)).astype(np.float64) * out).sum())


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


def eval_bound(e_ref, g_ref):
    """The evaluation's own bound on shared maps (north star: 1e-4 kcal/mol absolute / 1e-3
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


def check_best_chains(ctx, grid, lig, lid, e, conf, heavy, k_best=8):
    """Property checks of one ligand's docking output: every value finite; the energies of the k
    best chains are the oracle's evaluation (search caps: v = 10, slope 1e6) of the returned
    conformations on the same maps; the heavy-atom coordinates are those of the returned
    conformation."""
    assert np.isfinite(e).all() and np.isfinite(conf).all() and np.isfinite(heavy).all()
    worst = 0.0
    for i in np.argsort(e)[:k_best]:
        c = conf[i].astype(np.float64)
        e_ref = orc.eval_deriv(grid, lig, c)[0]
        tol = max(1e-3, 1e-3 * abs(e_ref))
        assert abs(e[i] - e_ref) <= tol, (lid, i, e[i], e_ref)
        worst = max(worst, abs(e[i] - e_ref))
        x = lig.set_conf(c)[lig.atoms()["el"] != 0]
        assert np.abs(x - heavy[i]).max() < 2e-3, (lid, i)
    return worst
