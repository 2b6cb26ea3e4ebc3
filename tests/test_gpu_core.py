"""GPU parity tests: maps, energy/gradient and docking through the C ABI vs the FP64 oracle.

Tolerances (stated per test):
  maps      |gpu - ref| <= 1e-4 * max(1, |ref|) + slack, where slack is the oracle's bound on the
            change caused by an r^2 within 0.003 LUT bins of a bin edge (FP32 vs FP64 rounding of
            r^2 is below 1e-3 bins here and can move such an r^2 to the neighbouring bin).
  energies  |gpu - ref| <= max(1e-4, 1e-3 * |ref|) kcal/mol (the parity target).
  gradients |gpu - ref| <= 1e-3 * max(1, |ref|_inf) + 1e-4 per component.
"""
import numpy as np
import pytest

import oracle as orc
from conftest import golden

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_receptor(n_atoms=2000, cube=36.0, cavity=8.0, seed=42):
    import vdock
    return vdock.synth_receptor(n_atoms, cube, cavity, seed)


@pytest.fixture(scope="module")
def lut():
    return orc.Lut()


@pytest.fixture(scope="module")
def synth_setup(ctx, lut):
    """Config-0 style receptor (2000 atoms, 36 A cube, 8 A cavity) with a 12 A box."""
    import vdock
    xyz, xs = make_receptor()
    box = vdock.box_from_center([0, 0, 0], [12, 12, 12], 0.375)
    ctx.set_receptor(xyz, xs)
    say("grid populate synthetic 12A box")
    ms = ctx.grid_populate(box)
    maps = ctx.grid_download()
    b, e, n = box.as_tuple()
    g = orc.Grid(lut, b, e, n)
    g.set_receptor(xyz.astype(np.float64), xs)
    slack = g.populate_with_slack(0.003)
    ref = g.maps()
    return dict(box=box, maps=maps, ref=ref, slack=slack, grid=g, ms=ms, xyz=xyz, xs=xs)


def check_maps(maps, ref, slack):
    tol = 1e-4 * np.maximum(1.0, np.abs(ref)) + slack
    bad = np.abs(maps - ref) > tol
    strict = slack == 0
    return int(bad.sum()), float(strict.mean())


def test_grid_maps_synthetic(synth_setup):
    s = synth_setup
    nbad, frac_strict = check_maps(s["maps"], s["ref"], s["slack"])
    say("maps", s["maps"].shape, "bad", nbad, "strict fraction", frac_strict, "ms", s["ms"])
    assert nbad == 0
    assert frac_strict > 0.3


def test_grid_maps_1u1b(ctx, lut):
    """Real receptor (1u1b, reference example) and the reference's 1u1b box."""
    import vdock
    text = golden("1u1b_protein.pdbqt")
    n = ctx.set_receptor_pdbqt(text)
    xyz_o, xs_o, _ = orc.parse_receptor(text)
    xyz_g, xs_g = ctx.receptor()
    assert n == len(xs_o) and np.array_equal(xs_g, xs_o)
    # V11/input_file_example/1u1b_config.txt, shrunk to 12 A for the FP64 oracle's run time
    box = vdock.box_from_center([-30.661, -29.540, 17.010], [12.0, 12.0, 12.0], 0.375)
    say("grid populate 1u1b")
    ctx.grid_populate(box)
    maps = ctx.grid_download()
    b, e, nn = box.as_tuple()
    g = orc.Grid(lut, b, e, nn)
    g.set_receptor(xyz_o, xs_o)
    slack = g.populate_with_slack(0.003)
    nbad, frac = check_maps(maps, g.maps(), slack)
    say("1u1b maps bad", nbad, "strict", frac)
    assert nbad == 0


def random_confs(rng, n, T, center, spread):
    conf = np.zeros((n, 7 + T))
    conf[:, :3] = np.asarray(center) + rng.uniform(-spread, spread, (n, 3))
    q = rng.normal(size=(n, 4))
    conf[:, 3:7] = q / np.linalg.norm(q, axis=1, keepdims=True)
    conf[:, 7:] = rng.uniform(-np.pi, np.pi, (n, T))
    return conf


LIGANDS = [("1u1b", None), ("2bm2", None), ("syn20", (20, 5, 1)), ("syn32", (32, 8, 2)), ("syn12", (12, 2, 3))]


def ligand_text(name, spec):
    import vdock
    if spec is None:
        return golden(f"{name}_ligand.pdbqt")
    return vdock.synth_ligand(*spec)


@pytest.mark.parametrize("name,spec", LIGANDS)
@pytest.mark.parametrize("cpw", [1, 2, 4])
def test_eval_matches_oracle(ctx, synth_setup, name, spec, cpw):
    s = synth_setup
    ctx.set_receptor(s["xyz"], s["xs"])
    ctx.grid_upload(s["box"], s["maps"])
    text = ligand_text(name, spec)
    lig = orc.Ligand(text)
    if lig.T + 7 > 64 // cpw:
        pytest.skip("ligand needs a wider lane group")
    lid = ctx.add_ligand(text)
    rng = np.random.default_rng(7)
    conf = random_confs(rng, 48, lig.T, [0, 0, 0], 1.5)
    say("eval", name, "cpw", cpw)
    e_gpu, g_gpu = ctx.eval(lid, conf.astype(np.float32), chains_per_wave=cpw)
    g_on_gpu_maps = orc.Grid(s["grid"].lut, *s["box"].as_tuple())
    g_on_gpu_maps.set_maps(s["maps"].astype(np.float64))
    worst_e = worst_g = 0.0
    for i in range(len(conf)):
        # same maps (FP32 values) in FP64 arithmetic: isolates the evaluation
        e_ref, g_ref, _, _ = orc.eval_deriv(g_on_gpu_maps, lig, conf[i].astype(np.float32).astype(np.float64))
        tol_e = max(1e-4, 1e-3 * abs(e_ref))
        assert abs(e_gpu[i] - e_ref) <= tol_e, (i, e_gpu[i], e_ref)
        tol_g = 1e-3 * max(1.0, np.abs(g_ref).max()) + 1e-4
        assert np.all(np.abs(g_gpu[i] - g_ref) <= tol_g), (i, g_gpu[i], g_ref)
        worst_e = max(worst_e, abs(e_gpu[i] - e_ref) / max(1.0, abs(e_ref)))
        worst_g = max(worst_g, np.abs(g_gpu[i] - g_ref).max() / max(1.0, np.abs(g_ref).max()))
    say("worst rel e", worst_e, "worst rel g", worst_g)


def test_eval_vs_fp64_maps(ctx, synth_setup):
    """End to end against the FP64 oracle (its own FP64 maps): error stated, bound as above."""
    s = synth_setup
    ctx.set_receptor(s["xyz"], s["xs"])
    ctx.grid_upload(s["box"], s["maps"])
    text = ligand_text("syn32", (32, 8, 2))
    lig = orc.Ligand(text)
    lid = ctx.add_ligand(text)
    conf = random_confs(np.random.default_rng(3), 64, lig.T, [0, 0, 0], 1.0)
    e_gpu, g_gpu = ctx.eval(lid, conf.astype(np.float32))
    errs, gerrs = [], []
    for i in range(len(conf)):
        e_ref, g_ref, _, _ = orc.eval_deriv(s["grid"], lig, conf[i].astype(np.float32).astype(np.float64))
        errs.append(abs(e_gpu[i] - e_ref))
        assert abs(e_gpu[i] - e_ref) <= max(1e-4, 1e-3 * abs(e_ref)), (i, e_gpu[i], e_ref)
        tol_g = 1e-3 * max(1.0, np.abs(g_ref).max()) + 1e-4
        gerrs.append(np.abs(g_gpu[i] - g_ref).max() / max(1.0, np.abs(g_ref).max()))
        assert np.all(np.abs(g_gpu[i] - g_ref) <= tol_g), (i, np.abs(g_gpu[i] - g_ref).max())
    say("max |e_gpu - e_fp64|", max(errs), "max relative gradient error", max(gerrs))


def test_dock_small(ctx, synth_setup):
    """A short docking run: finite results, energies consistent with a re-evaluation of the
    returned conformations, best pose better than the random starting points."""
    import vdock
    s = synth_setup
    ctx.set_receptor(s["xyz"], s["xs"])
    ctx.grid_upload(s["box"], s["maps"])
    text = ligand_text("syn20", (20, 5, 1))
    lid = ctx.add_ligand(text)
    p = vdock.default_params(chains=256, search_depth=4, bfgs_steps=12, seed=5)
    say("dock small")
    e, conf, heavy, st = ctx.dock(lid, p)
    say("best", e.min(), "evals", st.evals, "ms", st.kernel_ms)
    assert np.isfinite(e).all() and st.evals > 256
    e2, _ = ctx.eval(lid, conf)
    assert np.allclose(e2, e, atol=1e-3, rtol=1e-4)
    assert e.min() < 0.0


def test_batch_matches_plain(ctx, synth_setup):
    """Persistent batched launch and the per-ligand launch loop run the same chains, bit for bit:
    both run each ligand's chains at the same lanes per chain and register budget, and each chain
    draws from its ligand's own random stream."""
    import vdock
    s = synth_setup
    ctx.grid_upload(s["box"], s["maps"])
    ids = [ctx.add_ligand(vdock.synth_ligand(10 + 3 * k, k % 4, 100 + k)) for k in range(6)]
    p = vdock.default_params(chains=64, search_depth=3, bfgs_steps=8, seed=9, chains_per_wave=2)
    say("batch persistent")
    ctx.dock_batch(ids, p)
    best_b = ctx.batch_best()
    rb = [ctx.batch_results(k, p) for k in range(len(ids))]
    say("batch plain")
    ctx.dock_batch(ids, p, plain=True)
    best_p = ctx.batch_best()
    rp = [ctx.batch_results(k, p) for k in range(len(ids))]
    say("best batched", best_b, "plain", best_p)
    assert np.array_equal(best_b, best_p)
    for a, b in zip(rb, rp):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("policy", [0, 1, 2])
def test_class_overlap_same_results(ctx, synth_setup, policy):
    """The lanes-per-chain classes of a batch run at the same time on their own streams (the
    default) or one after the other: the same chains either way.  Each chain's trajectory depends
    only on its ligand, chain index and seed (histories are per chain), so the Vina and QuickVina2
    policies give the same results bit for bit; QuickVina-W's gate reads a buffer that all chains
    of a ligand fill in the order they run, so there only completion and finite results are
    checked.  With overlap the two launches' device intervals intersect."""
    import vdock
    s = synth_setup
    ctx.grid_upload(s["box"], s["maps"])
    ids = [ctx.add_ligand(vdock.synth_ligand(14 + 3 * k if k % 2 == 0 else 40 + k, 2 if k % 2 == 0 else 12, 300 + k))
           for k in range(8)]
    assert sorted({min(ctx.ligand_info(i).n_torsions, 10) for i in ids}) == [2, 10]
    p = vdock.default_params(chains=64, search_depth=3, bfgs_steps=8, seed=5, chains_per_wave=4, waves_per_simd=3,
                             policy=policy)
    out = {}
    for on in (False, True):
        ctx.set_class_overlap(on)
        say("class overlap", on, "policy", policy)
        st = ctx.dock_batch(ids, p)
        out[on] = ([ctx.batch_results(k, p) for k in range(len(ids))], ctx.last_launches(), st)
    ctx.set_class_overlap(True)
    launches = sorted(out[True][1], key=lambda r: r.start_ms)
    assert sorted(r.lanes_per_chain for r in launches) == [16, 32]
    a, b = launches
    assert b.start_ms < a.start_ms + a.kernel_ms
    for k in range(len(ids)):
        e, conf, heavy = out[True][0][k]
        assert np.isfinite(e).all() and np.isfinite(conf).all() and np.isfinite(heavy).all()
        if policy != 2:
            for x, y in zip(out[False][0][k], out[True][0][k]):
                assert np.array_equal(x, y)
    if policy != 2:
        assert out[True][2].evals == out[False][2].evals


def test_class_overlap_on_a_later_context(ctx, synth_setup):
    """A context made while others exist still runs the two classes of a batch at the same time:
    the contexts of a process share the library's streams, so its overlapped launches do not land
    on streams that share a hardware queue (which would run them one after the other)."""
    import vdock
    s = synth_setup
    other = vdock.Context(0)
    try:
        other.grid_upload(s["box"], s["maps"])
        ids = [other.add_ligand(vdock.synth_ligand(14 if k % 2 == 0 else 40, 2 if k % 2 == 0 else 12, 400 + k))
               for k in range(6)]
        p = vdock.default_params(chains=256, search_depth=3, bfgs_steps=8, seed=3)
        other.dock_batch(ids, p)
        a, b = sorted(other.last_launches(), key=lambda r: r.start_ms)
        say("later context:", a.kernel, a.start_ms, a.kernel_ms, b.kernel, b.start_ms, b.kernel_ms)
        assert {a.lanes_per_chain, b.lanes_per_chain} == {16, 64}
        assert b.start_ms < a.start_ms + 0.5 * a.kernel_ms
        # the first context still works beside it
        ctx.grid_upload(s["box"], s["maps"])
        st = ctx.dock_batch(ids[:2], p)
        assert st.evals > 0
    finally:
        other.close()


def test_fixed_launch_shapes_and_autotune(ctx, synth_setup):
    """Left at 0, lanes per chain and waves per SIMD come from the fixed table per size bucket:
    16 lanes for a small ligand with few torsions, 64 for one with 12, and a second identical call
    gives the same chains bit for bit.  With timing-based tuning on, the
    batch still docks (trial launches, then the timed shapes) to finite results; turning it off
    again restores the table's shapes and results."""
    import vdock
    s = synth_setup
    ctx.grid_upload(s["box"], s["maps"])
    ids = [ctx.add_ligand(vdock.synth_ligand(14, 2, 401)), ctx.add_ligand(vdock.synth_ligand(40, 12, 402))]
    p = vdock.default_params(chains=64, search_depth=3, bfgs_steps=8, seed=9)
    runs = []
    for tune in (False, False, True, False):
        ctx.set_autotune(tune)
        say("autotune", tune)
        ctx.dock_batch(ids, p)
        runs.append(([ctx.batch_results(k, p) for k in range(len(ids))],
                     sorted((r.lanes_per_chain, r.waves_per_simd) for r in ctx.last_launches())))
    assert [g for g, _ in runs[0][1]] == [16, 64]
    for k in range(len(ids)):
        assert all(np.isfinite(x).all() for x in runs[2][0][k])
        for r in (runs[1], runs[3]):
            for x, y in zip(runs[0][0][k], r[0][k]):
                assert np.array_equal(x, y)
    assert runs[3][1] == runs[0][1]
