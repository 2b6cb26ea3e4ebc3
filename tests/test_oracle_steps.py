"""Step-level known answers for the oracle's optimisers and Monte Carlo step (CPU).

Each test derives its expected values by hand from the reference's rules, on problems small
enough to follow step by step, so that a plausible slip in the oracle (a curvature constant of
0.9, an initial step of 1, the textbook sign of the direction update, a temperature of 1, one
entity fewer in the mutation draw) fails at least one assertion.

  Lewis-Overton line search + RILC-BFGS   V11/inc/cuda/quasi_newton.cuh:1710-1806, 2172-2389
  dense BFGS + backtracking               V11/lib/bfgs.h:30-140
  Metropolis rule                         V11/lib/monte_carlo.cpp:35-39 (temperature 1.2)
  mutation                                V11/lib/mutate.cpp:62-127
"""
import numpy as np
import pytest

import oracle as orc


def test_line_search_known_answer_on_a_parabola():
    """f = x^2/2 from x0 = 10: g = 10, d = -10, first step 1/|d| = 0.1; dginit = -100, so the
    sufficient-decrease slope is 1e-4 x -100 and the curvature test is g.d >= 0.1 x -100 = -10.
    Trials at steps 0.1, 0.2, 0.4, 0.8 land at x = 9, 8, 6, 2 with g.d = -90, -80, -60, -20 < -10
    (the step doubles); step 1.6 lands at x = -6 with g.d = 60 >= -10 and f = 18 < 50 - 1.6e-4:
    accepted at count 5.  One iteration (max_steps 1) ends there."""
    f, x, pts, fs = orc.opt_trace(0, [10.0], [1.0], max_steps=1, mode=1)
    assert np.allclose(pts.ravel(), [10, 9, 8, 6, 2, -6], rtol=0, atol=1e-12)
    assert np.allclose(fs, [50, 40.5, 32, 18, 2, 18], rtol=0, atol=1e-12)
    assert x[0] == pytest.approx(-6.0, abs=1e-12) and f == pytest.approx(18.0, abs=1e-12)
    # the second iteration: s = 1.6 x -10 = -16, y = -6 - 10 = -16, the one-pair update gives
    # d = 6 (Newton's step on a parabola), tried at step 1: x = 0, gradient 0, the search ends
    f, x, pts, fs = orc.opt_trace(0, [10.0], [1.0], max_steps=50, mode=1)
    assert np.allclose(pts.ravel(), [10, 9, 8, 6, 2, -6, 0], rtol=0, atol=1e-12)
    assert x[0] == pytest.approx(0.0, abs=1e-12) and f == pytest.approx(0.0, abs=1e-12)


def test_line_search_brackets_and_bisects():
    """Every first trial moves the point by exactly 1 along -g (step 1/|g|).
    f = x^2/2 from 0.25: the trial at -0.75 (f 0.28125) fails sufficient decrease against
    f0 = 0.03125, so the step is bracketed and bisected: 2 -> x = -0.25, f = 0.03125, still not
    below f0 + 2 x 1e-4 x (-0.0625); 1 -> x = 0, accepted at count 3.
    f = 50 x^2 from 2 (g = 200): the trial at 1 has g.d = -20000, below the curvature threshold
    0.1 x -40000, so the step doubles: 0.01 -> x = 0, accepted at count 2."""
    _, x, pts, _ = orc.opt_trace(0, [0.25], [1.0], max_steps=1, mode=1)
    assert np.allclose(pts.ravel(), [0.25, -0.75, -0.25, 0.0], atol=1e-15)
    _, x, pts, _ = orc.opt_trace(0, [2.0], [100.0], max_steps=1, mode=1)
    assert np.allclose(pts.ravel(), [2.0, 1.0, 0.0], atol=1e-12)


def test_direction_update_keeps_the_reference_sign():
    """f = (x^2 + 10 y^2)/2 from (1, 1).  The first line search accepts its first trial, step
    1/sqrt(101) along -(1, 10): x1 = (0.90049628097900, 0.00496280979001).  Then s = x1 - x0,
    y = g1 - g0, alpha = s.(-g1)/y.s, and the reference forms d = (-g1 + alpha y)(y.s / y.y) +
    (alpha - beta) s with beta = y.d'/y.s (quasi_newton.cuh:2342-2350); the first trial of the
    second line search (step 1) lands at (0.81081222453975, -0.00810812224540).  The textbook
    two-loop recursion (-alpha y) would put it at (0.80830078830279, -0.00808300788303).  The
    second search then doubles three times and accepts (0.18302382946498, -0.09960464649326)."""
    _, _, pts, _ = orc.opt_trace(0, [1.0, 1.0], [1.0, 10.0], max_steps=2, mode=1)
    assert np.allclose(pts[1], [0.90049628097900, 0.00496280979001], atol=1e-12)
    assert np.allclose(pts[2], [0.81081222453975, -0.00810812224540], atol=1e-12)
    assert not np.allclose(pts[2], [0.80830078830279, -0.00808300788303], atol=1e-6)
    assert np.allclose(pts[5], [0.18302382946498, -0.09960464649326], atol=1e-12)
    assert len(pts) == 6


def test_dense_bfgs_backtracking_and_old_gradient_stop():
    """Dense BFGS (bfgs.h): H = I, p = -H g.  On f = 5 x^2 from 1: p = -10; alpha = 1, 1/2, 1/4
    land at -9, -4, -1.5, all above f0 + 1e-4 alpha p.g; alpha = 1/8 lands at -0.25 (f 0.3125):
    accepted.  The first step rescales H's diagonal to alpha y.p / y.y = 0.1 (y = -12.5), the exact
    inverse curvature, so the next step lands at 0.  The stop test looks at the gradient before
    the step (bfgs.h:90), so one more iteration runs at x = 0 with p = 0: ten trials, all at 0,
    none of them a strict decrease, then the loop ends: 1 + 4 + 1 + 10 evaluations."""
    f, x, pts, _ = orc.opt_trace(0, [1.0], [10.0], max_steps=300, mode=0)
    assert np.allclose(pts.ravel()[:6], [1, -9, -4, -1.5, -0.25, 0], atol=1e-12)
    assert np.allclose(pts.ravel()[6:], 0, atol=1e-15)
    assert len(pts) == 16
    assert f == pytest.approx(0.0, abs=1e-15) and x[0] == pytest.approx(0.0, abs=1e-15)


@pytest.fixture(scope="module")
def mc_scene():
    import vdock
    xyz, xs = vdock.synth_receptor(2000, 36.0, 8.0, 42)
    box = vdock.box_from_center([0, 0, 0], [16, 16, 16], 0.375)
    b, e, n = box.as_tuple()
    g = orc.Grid(orc.Lut(), b, e, n)
    g.set_receptor(xyz.astype(np.float64), xs)
    g.populate(True)
    text = vdock.synth_ligand(16, 4, 21)
    return dict(grid=g, lig=orc.Ligand(text), b=b, e=e)


def _traces(s, chains, depth, quirk=0, seed=3):
    p = orc.mc_params(chains, depth, 4, s["b"], s["e"], seed=seed, stream=5, quirk=quirk)
    return [orc.mc_trace(s["grid"], s["lig"], p, c) for c in range(chains)]


def test_metropolis_acceptance_frequencies(mc_scene):
    """After step 0 (always accepted), a candidate at least as good as the current conformation
    is kept, and a worse one with probability exp(-(e_cand - e_prev)/1.2).  Over every step of 384
    chains: the number of uphill moves kept must match the sum of those probabilities (a z score
    below 4, binomial variance), and in each range of the energy rise separately; with a
    temperature of 1 (or 1.44) the expected count would be several standard deviations away."""
    tr = _traces(mc_scene, 384, 60)
    dE, acc = [], []
    for t in tr:
        assert t["accept"][0]
        d = t["e_cand"][1:] - t["e_prev"][1:]
        dE.append(d)
        acc.append(t["accept"][1:])
    dE, acc = np.concatenate(dE), np.concatenate(acc)
    assert acc[dE < 0].all()
    up = dE > 0
    assert up.sum() > 5000
    for T_alt in (1.0, 1.44):       # the test can tell these apart from the reference's 1.2
        p_alt = np.exp(-dE[up] / T_alt)
        z_alt = (acc[up].sum() - p_alt.sum()) / np.sqrt((p_alt * (1 - p_alt)).sum())
        assert abs(z_alt) > 4, (T_alt, z_alt)
    p = np.exp(-dE[up] / 1.2)
    z = (acc[up].sum() - p.sum()) / np.sqrt((p * (1 - p)).sum())
    assert abs(z) < 4, z
    for lo, hi in ((0, 0.5), (0.5, 1.5), (1.5, 4.0)):
        m = up & (dE >= lo) & (dE < hi)
        q = np.exp(-dE[m] / 1.2)
        if m.sum() >= 50:
            assert abs(acc[m].sum() - q.sum()) < 4 * np.sqrt((q * (1 - q)).sum()) + 1, (lo, hi)


def test_mutation_draw_and_moves(mc_scene):
    """The mutation picks one of 2 + T entities uniformly (position, orientation, each torsion;
    mutate.cpp:62-68) -- chi-square against uniform -- and moves only that entity: the position
    by amplitude x a point of the unit ball (less than 2 A), the orientation by a rotation of
    angle below amplitude / r_g, r_g the heavy atoms' RMS distance from the root origin of the
    current conformation, one torsion to a fresh value in [-pi, pi).  With the reference GPU
    tables' draw (mutation_quirk, main_procedure_cu.cpp:591) the last torsion is never picked."""
    from scipy.stats import chisquare
    s = mc_scene
    lig, T = s["lig"], s["lig"].T
    tr = _traces(s, 64, 40)
    which = np.concatenate([t["which"] for t in tr])
    counts = np.bincount(which, minlength=2 + T)
    assert len(counts) == 2 + T
    assert chisquare(counts).pvalue > 1e-3, counts
    heavy = lig.atoms()["el"] != 0
    for t in tr:
        for k in range(len(t["which"])):
            b, m, w = t["before"][k], t["mutated"][k], t["which"][k]
            pos_same = np.array_equal(b[:3], m[:3])
            rot_same = np.array_equal(b[3:7], m[3:7])
            tor_diff = np.nonzero(b[7:] != m[7:])[0]
            if w == 0:
                assert rot_same and len(tor_diff) == 0 and 0 < np.linalg.norm(m[:3] - b[:3]) < 2.0
            elif w == 1:
                assert pos_same and len(tor_diff) == 0
                X = lig.set_conf(b)[heavy]
                rg = np.sqrt(((X - b[:3]) ** 2).sum(-1).mean())
                assert t["rg"][k] == pytest.approx(rg, rel=1e-9)
                angle = 2 * np.arccos(min(1.0, abs(float(np.dot(b[3:7], m[3:7])))))
                assert angle < 2.0 / rg + 1e-9
            else:
                assert pos_same and rot_same and set(tor_diff.tolist()) <= {w - 2}
                assert -np.pi <= m[7 + w - 2] <= np.pi
    quirk = np.concatenate([t["which"] for t in _traces(s, 32, 40, quirk=1)])
    assert quirk.max() <= T and (quirk == T).sum() > 0
    assert chisquare(np.bincount(quirk, minlength=1 + T)).pvalue > 1e-3
