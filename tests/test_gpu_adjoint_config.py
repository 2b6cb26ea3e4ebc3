"""The adjoint microbenchmark's workload as a parity case: a 40-heavy-atom, 12-torsion synthetic
ligand on the single-ligand scene's maps (8,000-atom receptor, 30 A box, 81^3 points), 2^20
conformations drawn on the device (translation uniform in the box, orientation a normalised
4-Gaussian, torsions uniform in [-pi, pi)), each evaluated forward + adjoint through the device
pointer path of vd_eval.

  - every energy and gradient is finite;
  - on a sample of the conformations, against the FP64 oracle's eval_deriv on the same FP32 maps
    (isolates the evaluation): |de| <= max(1e-4, 1e-3 |e|) and |dg| <= 1e-3 max(1, |g|_inf) + 1e-4
    per component, plus the FP32 rounding of the out-of-box penalty (slope x 8 ulp of the
    coordinate per axis on which a scored atom lies outside: box-uniform translations put many
    atoms outside);
  - against the oracle end to end (its own FP64 maps): the same bounds plus what the maps'
    differences can add, and the largest error is stated;
  - the host-pointer path gives the same numbers bit for bit.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import eval_bound, map_error_bounds, out_of_box_bound, say

pytestmark = pytest.mark.gpu


def test_adjoint_workload_matches_oracle(ctx):
    import vdock
    from vdock import workloads as W
    say("adjoint workload: single-ligand scene")
    xyz, xs, box = W.single_ligand_scene(vdock, ctx)
    assert tuple(box.dims) == (81, 81, 81)
    maps = ctx.grid_download()
    text = vdock.synth_ligand(40, 12, 777)
    lid = ctx.add_ligand(text)
    lig = orc.Ligand(text)
    assert lig.T == 12 and lig.n_heavy == 40
    T = lig.T
    n = 1 << 20
    d_conf, d_e, d_g = ctx.alloc(n * (7 + T) * 4), ctx.alloc(n * 4), ctx.alloc(n * (6 + T) * 4)
    try:
        ctx.random_confs(lid, n, 1, d_conf)
        say("eval 2^20 conformations (device pointers)")
        ms = ctx.eval_device(lid, d_conf, d_e, d_g, n)
        conf = ctx.to_host(d_conf, (n, 7 + T))
        e = ctx.to_host(d_e, (n,))
        g = ctx.to_host(d_g, (n, 6 + T))
    finally:
        for p in (d_conf, d_e, d_g):
            ctx.free(p)
    say("kernel ms", ms)
    assert np.isfinite(e).all() and np.isfinite(g).all()
    b, en, _ = box.as_tuple()
    assert (conf[:, :3] >= np.asarray(b) - 1e-4).all() and (conf[:, :3] <= np.asarray(en) + 1e-4).all()
    assert np.allclose(np.linalg.norm(conf[:, 3:7], axis=1), 1.0, atol=1e-5)
    assert (np.abs(conf[:, 7:]) <= np.pi + 1e-6).all()
    sample = np.random.default_rng(5).choice(n, 1024, replace=False)
    # the host-pointer path on the sample: identical numbers
    e_h, g_h = ctx.eval(lid, conf[sample])
    assert np.array_equal(e_h, e[sample]) and np.array_equal(g_h, g[sample])
    # against the oracle on the same FP32 maps
    shared = orc.Grid(orc.Lut(), *box.as_tuple())
    shared.set_maps(maps.astype(np.float64))
    fp64 = orc.Grid(orc.Lut(), *box.as_tuple())
    fp64.set_receptor(xyz.astype(np.float64), xs)
    slack = fp64.populate_with_slack(0.003)
    ref_maps = fp64.maps()
    maps_tol = 1e-4 * np.maximum(1.0, np.abs(ref_maps)) + slack
    assert not (np.abs(maps - ref_maps) > maps_tol).any()
    worst = {"shared": [0.0, 0.0], "fp64": [0.0, 0.0]}
    abs_err = 0.0
    for i in sample:
        c = conf[i].astype(np.float64)
        oob = out_of_box_bound(lig, box, c)
        for label, grid in (("shared", shared), ("fp64", fp64)):
            e_ref, g_ref, _, _ = orc.eval_deriv(grid, lig, c)
            tol_e, tol_g = eval_bound(e_ref, g_ref)
            tol_e += oob
            if label == "fp64":
                eb, gb = map_error_bounds(lig, box, maps_tol, c)
                tol_e += eb
                tol_g += gb
                abs_err = max(abs_err, abs(float(e[i]) - e_ref))
            assert abs(e[i] - e_ref) <= tol_e, (label, i, e[i], e_ref, tol_e)
            assert np.all(np.abs(g[i] - g_ref) <= tol_g), (label, i, np.abs(g[i] - g_ref).max(), tol_g)
            worst[label][0] = max(worst[label][0], abs(e[i] - e_ref) / tol_e)
            worst[label][1] = max(worst[label][1], float(np.abs(g[i] - g_ref).max()) / tol_g)
    say("worst |de|/bound, |dg|/bound: shared maps", worst["shared"], "FP64 oracle maps", worst["fp64"],
        "largest |e - e_fp64|", abs_err)
