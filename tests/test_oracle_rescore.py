"""Pins of the oracle's post-search step: refinement against the receptor atoms (non_cache),
rescoring (eval_adjusted + conf_independent) and the rmsd bounds (CPU only).

The anchor is the reference's own docking output for its example, V11/docking_result/
1u1b_docking_result.pdbqt (fixture tests/golden/1u1b_docking_result.pdbqt): nine refined poses
with the REMARK VINA RESULT score and rmsd lower/upper bounds the reference computed for each
(V11/main/main.cpp:440-470).  Rescoring its coordinates must give back those numbers; its poses
must be stationary points of the refinement objective.
"""
import numpy as np
import pytest

import oracle as orc
from conftest import golden

CENTER = [-30.661, -29.540, 17.010]   # V11/input_file_example/1u1b_config.txt
SIZE = [24.0, 26.25, 22.5]


def parse_models(text):
    """MODEL blocks of a Vina output file -> ([{serial: xyz}], [(score, lb, ub)])."""
    models, remarks, cur = [], [], None
    for ln in text.splitlines():
        if ln.startswith("MODEL"):
            cur = {}
        elif ln.startswith("REMARK VINA RESULT:"):
            remarks.append(tuple(float(x) for x in ln.split()[3:6]))
        elif ln.startswith(("ATOM", "HETATM")) and cur is not None:
            cur[int(ln[6:11])] = [float(ln[30:38]), float(ln[38:46]), float(ln[46:54])]
        elif ln.startswith("ENDMDL"):
            models.append(cur)
            cur = None
    return models, remarks


@pytest.fixture(scope="module")
def site():
    lut = orc.Lut()
    xyz, xs = orc.parse_receptor(golden("1u1b_protein.pdbqt"))[:2]
    b, e, n = orc.grid_dims(CENTER, SIZE)
    g = orc.Grid(lut, b, e, n)
    g.set_receptor(xyz, xs)
    lig = orc.Ligand(golden("1u1b_ligand.pdbqt"))
    models, remarks = parse_models(golden("1u1b_docking_result.pdbqt"))
    ser = lig.serials()
    X = [np.array([m[s] for s in ser]) for m in models]
    return g, lig, X, remarks


def test_rescoring_reproduces_reference_1u1b_results(site):
    """eval_adjusted (model.cpp:1428-1431) with best-mode intramolecular energy, num_tors_div
    (everything.cpp:296-303, terms.cpp:97-125) and rmsd_lower/upper_bound (model.cpp:1453-1521),
    on the reference's written coordinates (3 decimals): scores within the 0.05 of the printed
    rounding, rmsd bounds within the coordinate rounding."""
    g, lig, X, remarks = site
    assert len(X) == 9 and lig.n_atoms == 55
    assert lig.num_tors() == 15.0
    best_intra = orc.intra_fast(g, lig, X[0])
    for x, (score, lb_ref, ub_ref) in zip(X, remarks):
        assert orc.within(g, lig, x)
        e = orc.noncache_eval(g, lig, x) + orc.intra_fast(g, lig, x) - best_intra
        s = lig.conf_independent(e)
        assert abs(s - score) <= 0.05 + 1e-3, (s, score)
        lb, ub = lig.rmsd_bounds(x, X[0])
        assert lb == pytest.approx(lb_ref, abs=2e-3) and ub == pytest.approx(ub_ref, abs=2e-3)
        assert lb <= ub + 1e-12


def test_rescoring_catches_planted_mistakes(site):
    """The fixture separates the plausible slips: counting every rotor instead of num_tors' rule,
    the smooth instead of the fast table, or dropping the best-mode intramolecular energy."""
    g, lig, X, remarks = site
    best_intra = orc.intra_fast(g, lig, X[0])
    x, score = X[1], remarks[1][0]
    e_ok = orc.noncache_eval(g, lig, x) + orc.intra_fast(g, lig, x) - best_intra
    assert abs(lig.conf_independent(e_ok) - score) <= 0.051
    assert abs(e_ok / (1 + orc.VINA_WEIGHTS_ROT * lig.T) - score) > 0.06            # all T rotors
    assert abs(lig.conf_independent(e_ok + best_intra) - score) > 0.06              # no intra offset
    e_smooth = orc.noncache_eval(g, lig, x, deriv=True)[0] + orc.intra_fast(g, lig, x) - best_intra
    assert abs(e_smooth - e_ok) > 1e-4                                              # tables differ


def test_rescoring_pins_receptor_typing(site):
    """The printed scores depend on the XS type of every receptor atom within 8 A of the poses,
    so they pin the oracle's receptor typing (model::assign_types and bonded_to_HD /
    bonded_to_heteroatom, V11/lib/model.cpp:716-797; the XS classes of V11/lib/atom_constants.h:
    66-83, 249-270) for the classes around the pocket: each plausible slip
    -- the carbon polarity rule either way, a missed N-H donor, an O donor/acceptor mix-up --
    moves several of the nine scores beyond the print's rounding.  Classes with a handful of atoms
    far from the pocket (N_P, N_A, S_P here) are not reached by these poses."""
    g, lig, X, remarks = site
    xyz, xs = orc.parse_receptor(golden("1u1b_protein.pdbqt"))[:2]
    ref = np.array([r[0] for r in remarks])
    C_H, C_P, N_P, N_D, O_A, O_DA = 0, 1, 2, 3, 8, 9

    def worst(xs2, min_off):
        g2 = orc.Grid(g.lut, g.begin, g.end, g.n)
        g2.set_receptor(xyz, xs2)
        bi = orc.intra_fast(g2, lig, X[0])
        s = np.array([lig.conf_independent(orc.noncache_eval(g2, lig, x) + orc.intra_fast(g2, lig, x) - bi)
                      for x in X])
        d = np.abs(s - ref)
        return d.max(), int((d > 0.06).sum()) >= min_off

    assert worst(xs, 1)[0] <= 0.051
    for a, b, n in [(C_P, C_H, 3), (C_H, C_P, 2), (N_D, N_P, 9), (O_DA, O_A, 5), (O_A, O_DA, 5)]:
        xs2 = xs.copy()
        xs2[xs == a] = b
        assert (xs == a).sum() > 0
        assert worst(xs2, n)[1], (a, b)


def kabsch(P, Q):
    pc, qc = P.mean(0), Q.mean(0)
    U, _, Vt = np.linalg.svd((P - pc).T @ (Q - qc))
    D = np.diag([1.0, 1.0, np.sign(np.linalg.det(Vt.T @ U.T))])
    R = Vt.T @ D @ U.T
    return R, qc - R @ pc


def rot_to_quat(R):
    w = np.sqrt(max(0.0, 1 + R[0, 0] + R[1, 1] + R[2, 2])) / 2
    x = np.copysign(np.sqrt(max(0.0, 1 + R[0, 0] - R[1, 1] - R[2, 2])) / 2, R[2, 1] - R[1, 2])
    y = np.copysign(np.sqrt(max(0.0, 1 - R[0, 0] + R[1, 1] - R[2, 2])) / 2, R[0, 2] - R[2, 0])
    z = np.copysign(np.sqrt(max(0.0, 1 - R[0, 0] - R[1, 1] + R[2, 2])) / 2, R[1, 0] - R[0, 1])
    q = np.array([w, x, y, z])
    return q / np.linalg.norm(q)


def conf_from_coords(lig, X):
    """Position, orientation and torsions that put the ligand's atoms on X: the root by a Kabsch
    fit, then each torsion in preorder by a scan and a ternary search."""
    at, nd = lig.atoms(), lig.nodes()
    node = at["node"]
    c = lig.identity_conf()
    r0 = node == 0
    R, t = kabsch(at["coords0"][r0] - nd["origin0"][0], X[r0])
    c[:3], c[3:7] = t, rot_to_quat(R)
    for k in range(1, lig.n_nodes):
        m = node == k

        def err(a):
            cc = c.copy()
            cc[6 + k] = a
            return ((lig.set_conf(cc)[m] - X[m]) ** 2).sum()

        scan = np.linspace(-np.pi, np.pi, 721)
        a = scan[np.argmin([err(s) for s in scan])]
        lo, hi = a - 0.01, a + 0.01
        for _ in range(60):
            m1, m2 = lo + (hi - lo) / 3, hi - (hi - lo) / 3
            if err(m1) < err(m2):
                hi = m2
            else:
                lo = m1
        c[6 + k] = (lo + hi) / 2
    return c


@pytest.fixture(scope="module")
def golden_confs(site):
    g, lig, X, _ = site
    confs = [conf_from_coords(lig, x) for x in X]
    for c, x in zip(confs, X):
        assert np.sqrt(((lig.set_conf(c) - x) ** 2).sum(1).mean()) < 5e-3
    return np.array(confs)


def test_reference_poses_are_refinement_fixed_points(site, golden_confs):
    """refine_structure (main.cpp:139-157) started from the reference's refined poses: they are
    near-stationary points of its objective (non_cache inter + intra, v = 1000), so the
    refinement moves them by a small fraction of an angstrom, never raises the energy and keeps
    them in the box; a torsional perturbation of the same pose has a gradient 50x larger."""
    g, lig, X, _ = site
    steps = orc.refine_steps(lig)
    assert steps == (30 + 55) // 3
    rng = np.random.default_rng(0)
    for c in golden_confs:
        e0, gr = orc.noncache_eval_deriv(g, lig, c)
        er, cr, xr = orc.refine(g, lig, c, steps)
        assert er <= e0 + 1e-9 and e0 - er < 0.3
        assert np.sqrt(((xr - lig.set_conf(c)) ** 2).sum(1).mean()) < 0.2
        assert orc.within(g, lig, xr)
        cp = c.copy()
        cp[7:] += rng.normal(0, 0.3, lig.T)
        assert np.linalg.norm(orc.noncache_eval_deriv(g, lig, cp)[1]) > 50 * np.linalg.norm(gr)


def test_postprocess_reproduces_reference_modes(site, golden_confs):
    """The whole per-ligand post-processing (main.cpp:425-470) run on the reference's nine poses
    gives back its nine printed scores (as rounded), best first, and rmsd bounds against the
    best mode close to the printed ones."""
    g, lig, X, remarks = site
    modes = orc.postprocess(g, lig, golden_confs, orc.refine_steps(lig))
    assert len(modes) == 9
    assert [round(m["e"], 1) for m in modes] == sorted(r[0] for r in remarks)
    assert modes[0]["lb"] == 0 and modes[0]["ub"] == 0
    assert all(a["e"] <= b["e"] for a, b in zip(modes, modes[1:]))
    ref_lb = sorted(r[1] for r in remarks)
    assert np.allclose(sorted(m["lb"] for m in modes), ref_lb, atol=0.08)


def test_noncache_box_penalty_and_gradient(site, golden_confs):
    """non_cache (non_cache.cpp:38-128): outside the box each typed atom is scored at its clamped
    position plus slope x its distance out; the conformation gradient matches FP64 central
    differences of the smooth energy (the tabulated dor bounds the agreement)."""
    g, lig, X, _ = site
    x = X[0]
    el = lig.atoms()["xs"]
    typed = el < orc.N_XS
    shift = np.array([g.end[0] - x[:, 0].min() + 1.5, 0, 0])   # every atom 1.5+ A beyond the +x face
    xo = x + shift
    clamped = xo.copy()
    clamped[:, 0] = g.end[0]
    slope = 7.0
    pen = slope * (xo[typed, 0] - g.end[0]).sum()
    e_out = orc.noncache_eval(g, lig, xo, slope=slope)
    e_clamped = orc.noncache_eval(g, lig, clamped, slope=slope)
    assert e_out == pytest.approx(e_clamped + pen, rel=1e-12)
    _, f = orc.noncache_eval(g, lig, xo, slope=slope, deriv=True)
    _, fc = orc.noncache_eval(g, lig, clamped, slope=slope, deriv=True)
    assert np.allclose(f[typed, 0] - fc[typed, 0], slope)
    assert np.allclose(f[~typed], 0)
    assert not orc.within(g, lig, xo)

    c = golden_confs[2].copy()
    c[0] += 0.37
    _, gr = orc.noncache_eval_deriv(g, lig, c)
    h = 1e-5
    fd = np.zeros_like(gr)
    for i in range(len(gr)):
        d = np.zeros(len(gr))
        d[i] = 1.0
        fd[i] = (orc.noncache_eval_deriv(g, lig, lig.increment(c, d, h))[0]
                 - orc.noncache_eval_deriv(g, lig, lig.increment(c, d, -h))[0]) / (2 * h)
    assert np.abs(gr - fd).max() <= 2e-2 * max(1.0, np.abs(fd).max())


def test_refine_pulls_a_pose_into_the_box(site, golden_confs):
    """A pose pushed half out of the box is refined with rising slopes until its heavy atoms are
    inside (then it scores a finite energy); a pose far outside that the step budget cannot bring
    back scores max_fl (main.cpp:153-155)."""
    g, lig, X, _ = site
    c = golden_confs[0].copy()
    c[0] += g.end[0] - lig.set_conf(c)[:, 0].max() + 1.0
    er, cr, xr = orc.refine(g, lig, c, orc.refine_steps(lig))
    assert np.isfinite(er) and er < 1e300 and orc.within(g, lig, xr)
    far = golden_confs[0].copy()
    far[0] += 80.0
    er, _, xr = orc.refine(g, lig, far, 1)
    assert er > 1e300 and not orc.within(g, lig, xr)
