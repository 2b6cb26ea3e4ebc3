"""The device path against the oracle's frozen outputs (tests/golden/oracle_*.npz): the maps of a
9^3 patch of the 1u1b receptor, and energy and gradient of 64 fixed conformations of the 1u1b
ligand (the reference's nine docked modes, perturbations of them, random poses) on maps the
device populates over the example's whole search box."""
import os

import numpy as np
import pytest

import oracle as orc
from conftest import GOLDEN, golden

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def box_of(fx):
    import vdock
    b, e, n = (np.asarray(fx[k]) for k in ("begin", "end", "n"))
    center, size = (b + e) / 2, e - b
    box = vdock.box_from_center(list(center), list(size), 0.375)
    bb, ee, nn = box.as_tuple()
    assert np.allclose(bb, b, atol=1e-9) and np.allclose(ee, e, atol=1e-9) and list(nn) == list(n)
    return box


def test_maps_patch_matches_frozen_oracle_maps(ctx):
    """Every map value within 1e-4 relative, plus the oracle's own slack where an atom sits
    within 0.003 A of a LUT bin edge or the cutoff (FP32 and FP64 distances may fall on either
    side there)."""
    fx = load("oracle_maps_1u1b.npz")
    ctx.set_receptor_pdbqt(golden("1u1b_protein.pdbqt"))
    box = box_of(fx)
    say("fixture maps patch", tuple(fx["n"]))
    ctx.grid_populate(box)
    maps = ctx.grid_download()
    ref = fx["maps"]
    assert maps.shape == ref.shape
    xyz, xs, _ = orc.parse_receptor(golden("1u1b_protein.pdbqt"))
    g = orc.Grid(orc.Lut(), *box.as_tuple())
    g.set_receptor(xyz, xs)
    slack = g.populate_with_slack(0.003)
    assert np.allclose(g.maps(), ref, rtol=1e-12, atol=1e-12)
    tol = 1e-4 * np.maximum(1.0, np.abs(ref)) + slack
    bad = np.abs(maps - ref) > tol
    say("patch max |d|", float(np.abs(maps - ref).max()), "points with slack", int((slack > 0).sum()))
    assert not bad.any()


@pytest.fixture(scope="module")
def whole_box(ctx):
    """The example's search box: the oracle's FP64 maps with their slack, the device's maps."""
    fx = load("oracle_eval_1u1b.npz")
    ctx.set_receptor_pdbqt(golden("1u1b_protein.pdbqt"))
    box = box_of(fx)
    say("fixture eval populate", tuple(fx["n"]))
    ctx.grid_populate(box)
    dev = ctx.grid_download()
    xyz, xs, _ = orc.parse_receptor(golden("1u1b_protein.pdbqt"))
    g = orc.Grid(orc.Lut(), *box.as_tuple())
    g.set_receptor(xyz, xs)
    slack = g.populate_with_slack(0.003)
    lig = orc.Ligand(golden("1u1b_ligand.pdbqt"))
    return dict(fx=fx, box=box, dev=dev, ref=g.maps(), slack=slack, lig=lig,
                lid=ctx.add_ligand(golden("1u1b_ligand.pdbqt")))


def map_error_bounds(w, tol, conf):
    """What map errors of at most `tol` (per map point) can do to a pose's energy and gradient:
    the energy moves by at most the trilinear interpolation of tol at each scored atom (curl
    only shrinks differences); an atom's force by at most sqrt(3)/granularity times the largest
    tol of its cell's corners; a gradient component -- a sum of forces, or of torques about an
    atom of the ligand -- by at most the sum of those times max(1, ligand diameter)."""
    lig, box = w["lig"], w["box"]
    b, e, n = (np.asarray(v, dtype=np.float64) for v in box.as_tuple())
    factor = n / (e - b)
    at = lig.atoms()
    scored = at["xs"] < orc.N_XS
    X = lig.set_conf(conf)
    diam = max(1.0, float(np.sqrt(((X[:, None] - X[None]) ** 2).sum(-1)).max()))
    eb = fb = 0.0
    for i in np.nonzero(scored)[0]:
        s = np.clip((X[i] - b) * factor, 0, n)
        a = np.minimum(np.floor(s).astype(int), n.astype(int) - 1)
        f = s - a
        c = tol[at["xs"][i], a[2]:a[2] + 2, a[1]:a[1] + 2, a[0]:a[0] + 2]
        wz, wy, wx = (np.array([1 - f[k], f[k]]) for k in (2, 1, 0))
        eb += float((c * wz[:, None, None] * wy[None, :, None] * wx[None, None, :]).sum())
        fb += np.sqrt(3) * factor.max() * float(c.max())
    return eb, fb * diam


def out_of_box_bound(w, conf, slope=1e6, ulps=8):
    """The out-of-box penalty is slope x distance, not curled: FP32 coordinates (a few
    roundings of the kinematics, `ulps` units in the last place) move it by slope x that much
    per axis on which a scored atom lies outside."""
    lig, box = w["lig"], w["box"]
    b, e, _ = (np.asarray(v, dtype=np.float64) for v in box.as_tuple())
    X = lig.set_conf(conf)[lig.atoms()["xs"] < orc.N_XS]
    out = (X < b) | (X > e)
    return float((slope * ulps * np.spacing(np.abs(X).astype(np.float32)).astype(np.float64) * out).sum())


def check_eval(ctx, w, maps_tol, label):
    """Device energies and gradients against the frozen FP64 ones: the evaluation's own bound
    (max(1e-4, 1e-3 |e|) on energy, 1e-3 max(1, max|g|) + 1e-4 on the gradient, as for the
    evaluation on shared maps) plus what the maps' differences and the out-of-box penalty can
    add, per pose."""
    fx = w["fx"]
    conf = fx["conf"].astype(np.float32)
    assert np.array_equal(conf.astype(np.float64), fx["conf"])
    say("fixture eval", label, conf.shape)
    e, g = ctx.eval(w["lid"], conf)
    worst_e = worst_g = 0.0
    for k in range(len(conf)):
        eb, gb = map_error_bounds(w, maps_tol, fx["conf"][k])
        e_ref, g_ref = fx["e"][k], fx["grad"][k]
        tol_e = max(1e-4, 1e-3 * abs(e_ref)) + eb + out_of_box_bound(w, fx["conf"][k])
        tol_g = 1e-3 * max(1.0, np.abs(g_ref).max()) + 1e-4 + gb
        assert abs(e[k] - e_ref) <= tol_e, (label, k, e[k], e_ref, tol_e)
        assert np.all(np.abs(g[k] - g_ref) <= tol_g), (label, k, np.abs(g[k] - g_ref).max(), tol_g)
        worst_e = max(worst_e, abs(e[k] - e_ref) / tol_e)
        worst_g = max(worst_g, np.abs(g[k] - g_ref).max() / tol_g)
    say(label, "worst |de|/bound", worst_e, "worst |dg|/bound", worst_g)


def test_eval_on_oracle_maps_matches_frozen_energies(ctx, whole_box):
    """The oracle's maps, uploaded as FP32: isolates the evaluation (the maps' only difference is
    their FP32 rounding)."""
    w = whole_box
    ctx.grid_upload(w["box"], w["ref"].astype(np.float32))
    check_eval(ctx, w, np.abs(w["ref"]) * 2.0 ** -24, "oracle maps")


def test_eval_on_device_maps_matches_frozen_energies(ctx, whole_box):
    """The device's own maps over the whole example box: first within 1e-4 relative plus the
    oracle's slack of the oracle's maps, then the evaluation on them with that map bound."""
    w = whole_box
    tol = 1e-4 * np.maximum(1.0, np.abs(w["ref"])) + w["slack"]
    d = np.abs(w["dev"] - w["ref"])
    say("whole box maps max |d|", float(d.max()), "points with slack", int((w["slack"] > 0).sum()))
    assert not (d > tol).any()
    ctx.grid_upload(w["box"], w["dev"])
    check_eval(ctx, w, tol + np.abs(w["ref"]) * 2.0 ** -24, "device maps")
