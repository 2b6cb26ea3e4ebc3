"""Grid maps and their corner records on boxes that do not fill whole 4x4x4 tiles.

Boxes of 1x1x1 steps (2 points per axis, a single cell), and anisotropic boxes of (1, 3, 6),
(5, 2, 9) and (7, 7, 3) steps, placed where the synthetic receptor's atoms reach them:
  - every map point against the FP64 oracle's maps (cache::populate, V11/lib/cache.cpp:148-208),
    tolerance as in test_gpu_core.py (1e-4 relative plus the oracle's bin-edge slack);
  - energy and gradient of a 20-atom, 5-torsion ligand at conformations in, around and outside the
    box, against the oracle's eval_deriv (model.cpp:1324-1334) on the same FP32 maps, after
    vd_grid_populate and again after vd_grid_upload of those maps: both read the corner records
    the library builds from the plain maps, so this checks their cell indexing on every shape.
    Tolerances: |e| within max(1e-4, 1e-3 |e_ref|), gradients within 1e-3 max(1, |g_ref|_inf) + 1e-4.
"""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (1, 3, 6), (5, 2, 9), (7, 7, 3)]
GRAN = 0.375
CENTER = [5.0, -3.0, 2.0]


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def scene(ctx):
    import vdock
    xyz, xs = vdock.synth_receptor(2000, 36.0, 8.0, 42)
    return dict(xyz=xyz, xs=xs, lut=orc.Lut())


def box_of(steps):
    import vdock
    # sizes a hair under whole steps, so ceil(size / granularity) gives exactly `steps`
    size = [GRAN * s - 1e-3 for s in steps]
    box = vdock.box_from_center(CENTER, size, GRAN)
    assert tuple(box.as_tuple()[2]) == tuple(steps)
    return box


def confs_around(rng, n, T, box):
    b, e, _ = box.as_tuple()
    lo, hi = np.asarray(b) - 2.0, np.asarray(e) + 2.0
    conf = np.zeros((n, 7 + T))
    conf[:, :3] = rng.uniform(lo, hi, (n, 3))
    q = rng.normal(size=(n, 4))
    conf[:, 3:7] = q / np.linalg.norm(q, axis=1, keepdims=True)
    conf[:, 7:] = rng.uniform(-np.pi, np.pi, (n, T))
    return conf


def check_eval(ctx, lid, lig, grid, conf):
    e_gpu, g_gpu = ctx.eval(lid, conf.astype(np.float32))
    for i in range(len(conf)):
        e_ref, g_ref, _, _ = orc.eval_deriv(grid, lig, conf[i].astype(np.float32).astype(np.float64))
        assert abs(e_gpu[i] - e_ref) <= max(1e-4, 1e-3 * abs(e_ref)), (i, e_gpu[i], e_ref)
        tol_g = 1e-3 * max(1.0, np.abs(g_ref).max()) + 1e-4
        assert np.all(np.abs(g_gpu[i] - g_ref) <= tol_g), (i, g_gpu[i], g_ref)


@pytest.mark.parametrize("steps", SHAPES)
def test_maps_and_corner_records(ctx, scene, steps):
    import vdock
    box = box_of(steps)
    say("grid shape", steps)
    ctx.set_receptor(scene["xyz"], scene["xs"])
    ctx.grid_populate(box)
    maps = ctx.grid_download()
    b, e, n = box.as_tuple()
    g = orc.Grid(scene["lut"], b, e, n)
    g.set_receptor(scene["xyz"].astype(np.float64), scene["xs"])
    slack = g.populate_with_slack(0.003)
    ref = g.maps()
    assert maps.shape == ref.shape
    bad = np.abs(maps - ref) > 1e-4 * np.maximum(1.0, np.abs(ref)) + slack
    say("points", ref[0].size, "bad", int(bad.sum()), "nonzero share", float((ref != 0).mean()))
    assert int(bad.sum()) == 0
    assert (ref != 0).mean() > 0.5          # the receptor reaches the box

    text = vdock.synth_ligand(20, 5, 1)
    lig = orc.Ligand(text)
    lid = ctx.add_ligand(text)
    on_gpu_maps = orc.Grid(scene["lut"], b, e, n)
    on_gpu_maps.set_maps(maps.astype(np.float64))
    conf = confs_around(np.random.default_rng(11 + sum(steps)), 24, lig.T, box)
    check_eval(ctx, lid, lig, on_gpu_maps, conf)
    # the same maps through the upload path
    ctx.grid_upload(box, maps)
    check_eval(ctx, lid, lig, on_gpu_maps, conf)
