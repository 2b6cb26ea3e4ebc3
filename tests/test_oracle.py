"""Pins of the FP64 oracle (CPU only).

The reference ships no golden numeric vectors for this path, so each oracle function is pinned
by known answers the reference's own files fix (the input coordinates of its example ligands),
published test vectors (Philox), closed forms, invariants and brute force -- chosen so that a
dropped term, a wrong sign or index, or a transposed operand fails one of them.
"""
import math

import numpy as np
import pytest

import oracle as orc
from conftest import golden

W = orc.VINA_WEIGHTS
XS_R = [1.9, 1.9, 1.8, 1.8, 1.8, 1.8, 1.7, 1.7, 1.7, 1.7, 2.0, 2.1, 1.5, 1.8, 2.0, 2.2, 1.2]
C_H, C_P, N_P, N_D, N_A, N_DA, O_P, O_D, O_A, O_DA = range(10)
F_H, Met_D = 12, 16


# ---------------------------------------------------------------- random stream
def test_philox_known_answers():
    """Random123 known-answer vectors for Philox4x32-10."""
    assert orc.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert orc.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert orc.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


# ---------------------------------------------------------------- scoring terms
def test_terms_closed_forms():
    # C_H-C_H at the optimal distance 3.8: gauss1 = 1, gauss2 = exp(-(3/2)^2), hydrophobic = 1
    e = orc.term_eval(C_H, C_H, 3.8)
    assert e == pytest.approx(W[0] + W[1] * math.exp(-2.25) + W[3], abs=1e-14)
    # repulsion switches on below the optimal distance only: d = -1 -> d^2 = 1
    r = 3.8 - 1.0
    e = orc.term_eval(C_P, C_P, r)
    assert e == pytest.approx(W[0] * math.exp(-4) + W[1] * math.exp(-4.0) + W[2] * 1.0, abs=1e-14)
    # hydrophobic ramp: d = 1.0 -> (1.0 - 1.5)/(0.5 - 1.5) = 0.5 ; C_P is not hydrophobic
    d = 1.0
    h = orc.term_eval(C_H, F_H, XS_R[C_H] + XS_R[F_H] + d) - orc.term_eval(C_P, F_H, XS_R[C_P] + XS_R[F_H] + d)
    assert h == pytest.approx(W[3] * 0.5, abs=1e-14)
    # hydrogen bond: donor N_D with acceptor O_A at d = -0.35 -> slope_step(0, -0.7, d) = 0.5;
    # donor-donor and acceptor-acceptor pairs get nothing; it is symmetric in the pair order
    rr = XS_R[N_D] + XS_R[O_A] - 0.35
    hb = orc.term_eval(N_D, O_A, rr) - orc.term_eval(N_P, O_P, XS_R[N_P] + XS_R[O_P] - 0.35)
    assert hb == pytest.approx(W[4] * 0.5, abs=1e-14)
    assert orc.term_eval(O_A, N_D, rr) == orc.term_eval(N_D, O_A, rr)
    assert orc.term_eval(N_D, N_D, rr) == pytest.approx(orc.term_eval(N_P, N_P, rr), abs=1e-15)
    # metal donor with acceptor gets the h-bond term, metal radius 1.2
    rm = XS_R[Met_D] + XS_R[O_A] - 1.0
    assert orc.term_eval(Met_D, O_A, rm) - orc.term_eval(Met_D, O_P, XS_R[Met_D] + XS_R[O_P] - 1.0) == \
        pytest.approx(W[4] * 1.0, abs=1e-14)


def test_lut_invariants():
    lut = orc.Lut()
    n = lut.n
    assert n == 2051
    rs = np.sqrt(np.arange(n) / 32.0)
    for t1, t2 in [(0, 0), (0, 8), (3, 8), (12, 16), (4, 9), (10, 11)]:
        tp = orc.tri_index(t1, t2)
        se, sd, fa = lut.tables(tp)
        for i in (0, 1, 100, 463, 1000, 2049, 2050):
            assert se[i] == pytest.approx(orc.term_eval(t1, t2, rs[i]), abs=1e-13)
        assert np.allclose(fa[:-1], 0.5 * (se[:-1] + se[1:]), atol=1e-14)
        assert fa[-1] == pytest.approx(0.5 * se[-1])
        mid = slice(1, n - 1)
        dor = (se[2:] - se[:-2]) / ((rs[2:] - rs[:-2]) * rs[mid])
        assert np.allclose(sd[mid], dor, rtol=1e-12, atol=1e-12)
        assert sd[0] == 0 and sd[-1] == 0
        # interpolation used by the intramolecular term: exact at bin centres, linear between
        e, d = lut.eval_deriv(tp, 10.0 / 32.0)
        assert e == pytest.approx(se[10]) and d == pytest.approx(sd[10])
        e, d = lut.eval_deriv(tp, 10.25 / 32.0)
        assert e == pytest.approx(0.75 * se[10] + 0.25 * se[11])
        assert lut.eval_fast(tp, 10.9 / 32.0) == fa[10]


# ---------------------------------------------------------------- maps
def small_receptor(n=300, cube=20.0, seed=1):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-cube / 2, cube / 2, (n, 3))
    xs = rng.integers(0, 18, n).astype(np.int32)   # 17 = untyped
    return xyz, xs


def test_cells_equal_brute_force():
    lut = orc.Lut()
    xyz, xs = small_receptor()
    b, e, n = orc.grid_dims([0.5, -0.3, 0.2], [7.0, 6.0, 8.0])
    g1 = orc.Grid(lut, b, e, n)
    g1.set_receptor(xyz, xs)
    g1.populate(True)
    g2 = orc.Grid(lut, b, e, n)
    g2.set_receptor(xyz, xs)
    g2.populate(False)
    assert np.array_equal(g1.maps(), g2.maps())


def test_single_atom_closed_form():
    lut = orc.Lut()
    a = np.array([[0.3, -0.2, 0.1]])
    for t1 in (C_H, O_A, 13):
        b, e, n = orc.grid_dims([0, 0, 0], [6.0, 6.0, 6.0])
        g = orc.Grid(lut, b, e, n)
        g.set_receptor(a, np.array([t1], dtype=np.int32))
        g.populate(True)
        maps = g.maps()
        nx, ny, nz = g.dims
        rng = np.random.default_rng(t1)
        for _ in range(50):
            x, y, z = rng.integers(0, nx), rng.integers(0, ny), rng.integers(0, nz)
            p = np.array(b) + np.array([x, y, z]) * 0.375
            r2 = float(((p - a[0]) ** 2).sum())
            for t2 in (0, 5, 8, 16):
                want = lut.tables(orc.tri_index(t1, t2))[2][int(32 * r2)] if r2 <= 64 else 0.0
                assert maps[t2, z, y, x] == pytest.approx(want, abs=1e-12)


def test_maps_translation_invariant():
    lut = orc.Lut()
    xyz, xs = small_receptor(200, 16.0, 3)
    b, e, n = orc.grid_dims([0, 0, 0], [5.0, 5.0, 5.0])
    g1 = orc.Grid(lut, b, e, n)
    g1.set_receptor(xyz, xs)
    s1 = g1.populate_with_slack(1e-6)
    shift = np.array([7.5, -3.0, 1.125])
    g2 = orc.Grid(lut, np.array(b) + shift, np.array(e) + shift, n)
    g2.set_receptor(xyz + shift, xs)
    g2.populate(True)
    diff = np.abs(g1.maps() - g2.maps())
    assert np.all(diff <= 1e-9 + s1)


def test_cell_culling_keeps_all_atoms_in_range():
    lut = orc.Lut()
    xyz, xs = small_receptor(400, 24.0, 5)
    b, e, n = orc.grid_dims([0, 0, 0], [9.0, 9.0, 9.0])
    g = orc.Grid(lut, b, e, n)
    g.set_receptor(xyz, xs)
    rng = np.random.default_rng(0)
    for _ in range(40):
        p = np.array(b) + rng.uniform(0, 1, 3) * (np.array(e) - np.array(b))
        cand = set(g.cell_candidates(p).tolist())
        inrange = {i for i in range(len(xs)) if xs[i] < 17 and ((xyz[i] - p) ** 2).sum() <= 64}
        assert inrange <= cand


def test_trilinear_closed_form():
    """Map filled with a linear function: interpolation is exact, gradient = slope, no curl for
    negative values; out of the box the penalty is slope * miss / factor."""
    lut = orc.Lut()
    b, e, n = orc.grid_dims([0, 0, 0], [3.0, 3.0, 3.0])
    g = orc.Grid(lut, b, e, n)
    nx, ny, nz = g.dims
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    px, py, pz = b[0] + 0.375 * xx, b[1] + 0.375 * yy, b[2] + 0.375 * zz
    lin = -5.0 + 0.3 * px - 0.2 * py + 0.1 * pz
    g.set_maps(np.broadcast_to(lin, (17,) + lin.shape).copy())
    p = [0.31, -0.52, 0.77]
    val, d = g.eval(4, p, slope=1e6, v=10.0)
    assert val == pytest.approx(-5.0 + 0.3 * p[0] - 0.2 * p[1] + 0.1 * p[2], abs=1e-12)
    assert np.allclose(d, [0.3, -0.2, 0.1], atol=1e-12)
    out = [e[0] + 0.5, 0.0, 0.0]
    val, d = g.eval(4, out, slope=1e6, v=10.0)
    edge = -5.0 + 0.3 * e[0]
    assert val == pytest.approx(edge + 1e6 * 0.5, rel=1e-12)
    assert d[0] == pytest.approx(1e6) and d[1] == pytest.approx(-0.2)


# ---------------------------------------------------------------- ligand model and kinematics
LIGS = ["1u1b", "2bm2", "2er7"]


@pytest.mark.parametrize("name", LIGS)
def test_identity_conformation_reproduces_input(name):
    """model::get_initial_conf (V11/lib/model.cpp:996-1003): the identity conformation gives
    back the PDBQT coordinates -- a known answer from the reference's example files."""
    lig = orc.Ligand(golden(f"{name}_ligand.pdbqt"))
    xyz = lig.set_conf(lig.identity_conf())
    assert np.abs(xyz - lig.atoms()["coords0"]).max() < 1e-12


def test_reference_ligand_sizes():
    """Torsion counts as the reference's files declare them (REMARK active torsions / TORSDOF)
    and as its MC log reports them for its examples."""
    lig = orc.Ligand(golden("1u1b_ligand.pdbqt"))
    assert (lig.n_atoms, lig.T, lig.torsdof) == (55, 16, 16)
    lig = orc.Ligand(golden("2bm2_ligand.pdbqt"))
    assert (lig.n_atoms, lig.T) == (33, 7)


@pytest.mark.parametrize("name", LIGS)
def test_rigid_bodies_and_axes_preserved(name):
    lig = orc.Ligand(golden(f"{name}_ligand.pdbqt"))
    at, nd = lig.atoms(), lig.nodes()
    rng = np.random.default_rng(11)
    c = lig.identity_conf()
    c[:3] += rng.normal(size=3)
    q = rng.normal(size=4)
    c[3:7] = q / np.linalg.norm(q)
    c[7:] = rng.uniform(-np.pi, np.pi, lig.T)
    xyz = lig.set_conf(c)
    x0 = at["coords0"]
    for k in range(lig.n_nodes):
        idx = np.arange(nd["begin"][k], nd["end"][k])
        d0 = np.linalg.norm(x0[idx, None] - x0[None, idx], axis=-1)
        d1 = np.linalg.norm(xyz[idx, None] - xyz[None, idx], axis=-1)
        assert np.abs(d0 - d1).max() < 1e-9
    # bonded atoms keep their bond length (bonds across rotatable axes included)
    for i, j in lig.bonds():
        assert abs(np.linalg.norm(xyz[i] - xyz[j]) - np.linalg.norm(x0[i] - x0[j])) < 1e-9
    # a torsion change leaves everything outside that node's subtree unchanged
    c2 = c.copy()
    k = lig.T  # last node in preorder is a leaf
    c2[7 + k - 1] += 0.7
    xyz2 = lig.set_conf(c2)
    moved = np.where(np.abs(xyz2 - xyz).max(axis=1) > 1e-12)[0]
    assert set(moved.tolist()) <= set(range(nd["begin"][k], nd["end"][k]))


def test_pairs_definition_brute_force():
    """Intramolecular pairs are exactly the typed atom pairs in different rigid frames that are
    more than three bonds apart (for a tree-shaped ligand, by graph distance)."""
    import vdock
    for seed in range(4):
        lig = orc.Ligand(vdock.synth_ligand(18, 4, seed))
        at, nd = lig.atoms(), lig.nodes()
        n = lig.n_atoms
        adj = [[] for _ in range(n)]
        for i, j in lig.bonds():
            adj[i].append(j)
            adj[j].append(i)
        pairs = {(int(a), int(b)) for _, a, b in lig.pairs()}
        # same rigid body: same node, or a node's axis atom (in the parent) with that node's atoms
        fixed = set()
        for k in range(lig.n_nodes):
            ids = list(range(nd["begin"][k], nd["end"][k]))
            for a in ids:
                for b in ids:
                    fixed.add((a, b))
        for i in range(n):
            dist = {i: 0}
            frontier = [i]
            while frontier:
                nxt = []
                for u in frontier:
                    for v in adj[u]:
                        if v not in dist:
                            dist[v] = dist[u] + 1
                            nxt.append(v)
                frontier = nxt
            for j in range(i + 1, n):
                if at["xs"][i] >= 17 or at["xs"][j] >= 17:
                    assert (i, j) not in pairs
                    continue
                if dist.get(j, 99) <= 3:
                    assert (i, j) not in pairs
                elif (i, j) not in fixed and at["node"][i] != at["node"][j]:
                    # far apart, different nodes: a pair unless both sit on one rotation axis frame
                    if (i, j) not in pairs:
                        pi, pj = at["node"][i], at["node"][j]
                        assert pi == nd["parent"][pj] or pj == nd["parent"][pi]


# ---------------------------------------------------------------- energy and gradient
@pytest.fixture(scope="module")
def scene():
    import vdock
    xyz, xs = vdock.synth_receptor(1500, 30.0, 6.0, 4)
    lut = orc.Lut()
    b, e, n = orc.grid_dims([0, 0, 0], [14.0, 14.0, 14.0])
    g = orc.Grid(lut, b, e, n)
    g.set_receptor(xyz.astype(np.float64), xs)
    g.populate(True)
    return g


def fd_grad(g, lig, c, h=1e-5):
    out = np.zeros(6 + lig.T)
    for i in range(6 + lig.T):
        d = np.zeros(6 + lig.T)
        d[i] = 1.0
        ep = orc.eval_deriv(g, lig, lig.increment(c, d, h))[0]
        em = orc.eval_deriv(g, lig, lig.increment(c, d, -h))[0]
        out[i] = (ep - em) / (2 * h)
    return out


@pytest.mark.parametrize("terms", ["inter", "intra"])
def test_gradient_matches_fp64_central_differences(scene, terms):
    """Adjoint (force, torque, torsion derivatives) vs FP64 central differences (h = 1e-5)
    taken through the conformation increment (conf.h:118-124).  The intramolecular gradient
    uses the tabulated dor (precalculate.h:121-140), which differs from the derivative of the
    interpolated energy by the table's discretisation error; its bound is looser."""
    import vdock
    orc.set_terms(inter=terms == "inter", intra=terms == "intra")
    try:
        rng = np.random.default_rng(5)
        ok = 0
        trials = 12
        for t in range(trials):
            lig = orc.Ligand(vdock.synth_ligand(16 + t, 4, 50 + t))
            c = lig.identity_conf()
            c[:3] = rng.uniform(-2, 2, 3)
            q = rng.normal(size=4)
            c[3:7] = q / np.linalg.norm(q)
            c[7:] = rng.uniform(-np.pi, np.pi, lig.T)
            e, g, _, _ = orc.eval_deriv(scene, lig, c)
            fd = fd_grad(scene, lig, c)
            scale = max(1.0, np.abs(fd).max())
            tol = 1e-6 if terms == "inter" else 2e-2
            if np.abs(g - fd).max() <= tol * scale:
                ok += 1
        # a few conformations can straddle a trilinear cell boundary within h
        assert ok >= trials - 1, ok
    finally:
        orc.set_terms(True, True)


# ---------------------------------------------------------------- optimisers
def test_dense_bfgs_quadratic_and_rosenbrock():
    coef = np.arange(1, 9, dtype=float)
    f, x, ne = orc.opt_test(0, np.ones(8), coef, max_steps=100, mode=0)
    assert f < 1e-10 and np.abs(x).max() < 1e-5
    f, x, ne = orc.opt_test(1, np.array([-1.2, 1.0]), max_steps=200, mode=0)
    assert f < 1e-8 and np.allclose(x, [1, 1], atol=1e-4)


def test_rilc_bfgs_on_closed_forms():
    """The reference one-pair update (sign as in quasi_newton.cuh:2342-2350) on quadratics:
    isotropic curvature is solved in one line search; anisotropic ones decrease monotonically
    and converge; it never ends above its start (revert rule)."""
    f, x, ne = orc.opt_test(0, np.array([3.0, -4.0, 1.0]), np.full(3, 2.0), max_steps=50, mode=1)
    assert f < 1e-10
    coef = np.array([1.0, 4.0, 9.0, 2.0])
    x0 = np.array([1.0, -1.0, 0.5, 2.0])
    f0 = 0.5 * (coef * x0 ** 2).sum()
    f, x, ne = orc.opt_test(0, x0, coef, max_steps=300, mode=1)
    assert f < 1e-6 * f0
    for steps in (1, 2, 5, 10):
        f_s, _, _ = orc.opt_test(1, np.array([-1.2, 1.0]), max_steps=steps, mode=1)
        assert f_s <= 24.2 + 1e-12


def test_local_search_improves_and_reverts(scene):
    import vdock
    lig = orc.Ligand(vdock.synth_ligand(20, 5, 9))
    rng = np.random.default_rng(1)
    for mode in (0, 1):
        c = lig.identity_conf()
        c[:3] = rng.uniform(-2, 2, 3)
        e0 = orc.eval_deriv(scene, lig, c)[0]
        e1, c1, ne = orc.local_search(scene, lig, c, 16, mode=mode)
        assert e1 <= e0 and ne >= 1
        assert e1 == pytest.approx(orc.eval_deriv(scene, lig, c1)[0], abs=1e-9)


def test_mc_chain_deterministic_and_consistent(scene):
    import vdock
    lig = orc.Ligand(vdock.synth_ligand(16, 3, 21))
    p = orc.mc_params(8, 4, 10, [-7, -7, -7], [7, 7, 7], seed=3, threads=2)
    e1, c1, h1, ev1 = orc.mc_run(scene, lig, p)
    e2, c2, h2, ev2 = orc.mc_run(scene, lig, p)
    assert np.array_equal(e1, e2) and np.array_equal(c1, c2) and ev1 == ev2
    for i in range(len(e1)):
        assert e1[i] == pytest.approx(orc.eval_deriv(scene, lig, c1[i])[0], abs=1e-9)
        xyz = lig.set_conf(c1[i])
        heavy = xyz[lig.atoms()["el"] != 0]
        assert np.abs(heavy - h1[i]).max() < 1e-9


def test_initial_conformations_in_box_and_normalised():
    import vdock
    lig = orc.Ligand(vdock.synth_ligand(12, 3, 2))
    p = orc.mc_params(64, 1, 1, [-3, -2, -1], [3, 2, 1], seed=1)
    cs = np.array([orc.initial_conf(lig, p, k) for k in range(64)])
    assert np.all(cs[:, :3] >= [-3, -2, -1]) and np.all(cs[:, :3] <= [3, 2, 1])
    assert np.allclose(np.linalg.norm(cs[:, 3:7], axis=1), 1.0)
    assert np.all(np.abs(cs[:, 7:]) <= np.pi)
    # uniform orientations: the mean rotation axis component is near 0
    assert np.abs(cs[:, 4:7].mean(axis=0)).max() < 0.2


def test_cluster_rules():
    """add_to_output_container (V11/lib/coords.cpp:47-62): a pose within min_rmsd of a kept one
    replaces it only if better; the container keeps the best max_size, sorted."""
    heavy = np.zeros((5, 2, 3))
    heavy[1] += 0.1        # close to pose 0
    heavy[2] += 5.0
    heavy[3] += 5.05       # close to pose 2
    heavy[4] += 10.0
    e = np.array([-5.0, -6.0, -4.0, -3.0, -1.0])
    kept = orc.cluster(e, heavy, 1.0, 2)
    assert list(kept) == [1, 2]
    kept = orc.cluster(e, heavy, 1.0, 20)
    assert list(kept) == [1, 2, 4]


# ---- QuickVina2 visited history (QV2/inc/cuda/visited.cuh:7-226, intended semantics) ----
# Points of a convex bowl f = sum_i (c_i - m_i)^2 over the change dimensions' coordinates
# (position x,y,z, orientation components a,b,c, torsions) with g_i = 2 (c_i - m_i).  On one
# side of the minimum in every dimension, each pair of points lies on one monotone slope, so
# every point explains every other; across the minimum no dimension agrees in sign.
VT = 2
VN = 6 + VT
VCAP = 10 * VN


def _bowl(rng, k, side, lo=0.2, hi=1.5, around=None, spread=0.05):
    m = np.linspace(-0.3, 0.3, VN)
    coords = (around + rng.uniform(-spread, spread, (k, VN))) if around is not None else \
        m + side * rng.uniform(lo, hi, (k, VN))
    x = np.zeros((k, 7 + VT))
    x[:, 0:6] = coords[:, 0:6]
    x[:, 6] = rng.uniform(-1, 1, k)                 # orientation d: in the distance only
    x[:, 7:] = coords[:, 6:]
    f = ((coords - m) ** 2).sum(axis=1)
    g = 2.0 * (coords - m)
    return x, f, g


def _cat(*parts):
    return tuple(np.concatenate([p[i] for p in parts]) for i in range(3))


def test_visited_not_full_is_always_interesting():
    rng = np.random.default_rng(5)
    x = rng.normal(size=(VCAP, 7 + VT)); f = rng.normal(size=VCAP); g = rng.normal(size=(VCAP, VN))
    assert orc.visited_feed(VT, x, f, g, mode=2).all()
    # the gate (mode 1) adds every interesting point, so the history fills after exactly VCAP
    assert orc.visited_feed(VT, x, f, g, mode=1).all()


def test_visited_convex_side_rule():
    rng = np.random.default_rng(6)
    right = _bowl(rng, VCAP, +1)
    more_right = _bowl(rng, 12, +1)
    left = _bowl(rng, 1, -1)
    a = orc.visited_feed(VT, *_cat(right, more_right), mode=2)
    assert a[:VCAP].all() and not a[VCAP:].any()
    b = orc.visited_feed(VT, *_cat(right, left), mode=2)
    assert b[-1]


def test_visited_ring_replaces_oldest():
    rng = np.random.default_rng(7)
    right, left = _bowl(rng, VCAP, +1), _bowl(rng, VCAP, -1)
    # after VCAP more adds every right-side entry has been overwritten by a left-side one
    assert not orc.visited_feed(VT, *_cat(right, left, _bowl(rng, 1, -1)), mode=2)[-1]
    assert orc.visited_feed(VT, *_cat(right, left, _bowl(rng, 1, +1)), mode=2)[-1]
    # one add short of that, the last right-side entry can still be among the neighbours
    part = tuple(a[: VCAP - 1] for a in left)
    q = _bowl(rng, 1, +1)
    assert orc.visited_feed(VT, *_cat(right, part, q), mode=2)[-1]


def test_visited_uses_exactly_n_nearest():
    rng = np.random.default_rng(8)
    centre = np.linspace(-0.3, 0.3, VN) + 1.0
    q = _bowl(rng, 1, +1, around=centre, spread=0.01)
    for near_count, expect in ((VN, False), (VN - 1, True)):
        near = _bowl(rng, near_count, +1, around=centre)        # explain the query
        far = _bowl(rng, VCAP - near_count, -1, lo=1.0, hi=2.0)  # never explain it, all farther
        hist = _cat(far, near)
        assert orc.visited_feed(VT, *_cat(hist, q), mode=2)[-1] == expect


def test_visited_order_of_history_does_not_matter():
    rng = np.random.default_rng(9)
    answers = []
    for trial in range(40):
        x = rng.normal(size=(VCAP + 1, 7 + VT)); f = rng.normal(size=VCAP + 1); g = rng.normal(size=(VCAP + 1, VN))
        a = orc.visited_feed(VT, x, f, g, mode=2)[-1]
        perm = np.concatenate([rng.permutation(VCAP), [VCAP]])
        b = orc.visited_feed(VT, x[perm], f[perm], g[perm], mode=2)[-1]
        assert a == b
        answers.append(a)
    assert 0 < sum(answers) < len(answers)          # both answers occur


# ---- QuickVina-W (QVW/inc/cuda/visited.cuh, quasi_newton.cuh:2180-2420, intended semantics) ----
# Same bowl points.  Individual buffers hold 5n entries; the global buffer is shared by the
# chains of a ligand, a ring of 8 entries per cell.
QIND = 5 * VN
QBOX = ([-4.0, -4.0, -4.0], [4.0, 4.0, 4.0])


def _qops(op, chain, pts):
    x, f, g = pts
    return [(op, chain, x[i], f[i], g[i]) for i in range(len(f))]


def _qfeed(ops, levels=0):
    return orc.qvw_feed(VT, ops, *QBOX, levels)


def test_qvw_levels_follow_get_n():
    """get_n (QVW visited.cuh): the smallest multiple of 3, i, with chains x depth < 2^(i+1);
    (i - 3) / 3 levels of halving per axis, at least 0, capped at 4 here."""
    assert orc.qvw_levels(256, 20) == 3          # 5120 < 8192 = 2^13
    assert orc.qvw_levels(8192, 30) == 4         # 245760 < 2^19 -> 5 levels, capped
    assert orc.qvw_levels(10, 10) == 1           # 100 < 128 = 2^7
    assert orc.qvw_levels(3, 5) == 0             # 15 < 16 = 2^4
    assert orc.qvw_levels(1, 1) == 0


def test_qvw_gate_needs_a_full_individual_buffer():
    rng = np.random.default_rng(21)
    right, q = _bowl(rng, QIND - 1, +1), _bowl(rng, 1, +1)
    ops = _qops(1, 0, right) + _qops(0, 0, q)
    assert _qfeed(ops)[-1] == 1                  # 5n - 1 explaining entries: still interesting
    ops = _qops(1, 0, _bowl(rng, QIND, +1)) + _qops(0, 0, q) + _qops(0, 0, _bowl(rng, 1, -1))
    assert _qfeed(ops)[-2:] == [0, 1]            # full: the same side is skipped, the other searched


def test_qvw_global_buffer_is_shared_and_consulted_first():
    rng = np.random.default_rng(22)
    full_right = _qops(1, 0, _bowl(rng, QIND, +1))
    q = _qops(0, 0, _bowl(rng, 1, +1))
    # another chain's search ended across the minimum: chain 0's point becomes interesting
    assert _qfeed(full_right + _qops(2, 1, _bowl(rng, 1, -1)) + q)[-1] == 1
    # same-side global entries explain it: the G-check examines them all, the individual
    # buffer the remaining 4n - 8 nearest, and the point is skipped
    out = _qfeed(full_right + _qops(2, 1, _bowl(rng, 8, +1)) + _qops(3, 0, _bowl(rng, 1, +1)) + q)
    assert out[-2] == min(8, VN) and out[-1] == 0


def test_qvw_global_ring_keeps_the_last_8():
    rng = np.random.default_rng(23)
    full_right = _qops(1, 0, _bowl(rng, QIND, +1))
    q = _qops(0, 0, _bowl(rng, 1, +1))
    left8 = _qops(2, 1, _bowl(rng, 8, -1))
    assert _qfeed(full_right + left8 + _qops(2, 1, _bowl(rng, 8, +1)) + q)[-1] == 0
    assert _qfeed(full_right + left8 + _qops(2, 1, _bowl(rng, 7, +1)) + q)[-1] == 1


def test_qvw_global_cells():
    """Cells of 8 A / 2^levels over the box, x fastest; outside the box clamps to the border cell;
    an entry in another cell is not consulted."""
    rng = np.random.default_rng(24)
    x, f, g = _bowl(rng, 2, -1)
    x[0, :3] = [-3.9, 0.1, 3.9]
    x[1, :3] = [10.0, -10.0, 0.0]
    out = _qfeed(_qops(2, 1, (x, f, g)), levels=2)
    assert out == [(3 * 4 + 2) * 4 + 0, (2 * 4 + 0) * 4 + 3]
    full_right = _qops(1, 0, _bowl(rng, QIND, +1))
    qx, qf, qg = _bowl(rng, 1, +1)
    qx[0, :3] = [-3.9, 0.1, 3.9]
    left = _bowl(rng, 1, -1)
    for pos, expect in (([-3.8, 0.2, 3.8], 1), ([3.8, 0.2, 3.8], 0)):
        lx = left[0].copy()
        lx[0, :3] = pos
        ops = full_right + _qops(2, 1, (lx, left[1], left[2])) + _qops(0, 0, (qx, qf, qg))
        assert _qfeed(ops, levels=2)[-1] == expect


def test_qvw_policy_runs_fewer_searches(scene):
    """The MC with the QuickVina-W gate skips searches once buffers fill: fewer evaluations than
    the Vina policy on the same chains, and a comparable best score."""
    import vdock
    lig = orc.Ligand(vdock.synth_ligand(12, 2, 9))
    b, e = [-6.0, -6.0, -6.0], [6.0, 6.0, 6.0]
    res = {}
    for pol in (0, 2):
        p = orc.mc_params(32, 96, 8, b, e, seed=3, threads=8, policy=pol)
        en, _, _, ev = orc.mc_run(scene, lig, p)
        res[pol] = (float(en.min()), ev)
    assert res[2][1] < 0.97 * res[0][1]
    assert res[2][0] <= res[0][0] + 0.5
