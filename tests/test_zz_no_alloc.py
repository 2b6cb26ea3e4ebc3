"""vd_dock_noalloc, the docking entry point that allocates no device memory (its buffers beyond
the caller's inputs and outputs live in a per-device scratch array), against vd_dock on the same
inputs: the same kernel and launch shape, so every chain's energy, conformation and heavy-atom
coordinates are equal bit for bit.  Two calls in a row, the second on a smaller ligand with fewer
chains (the array keeps the first call's contents), then the QuickVina2 policy (its histories in
the array too), given starting conformations, and a request beyond the array's size, which must be
refused without touching the outputs."""
import numpy as np
import pytest

from helpers import say

pytestmark = pytest.mark.gpu


def _noalloc(ctx, lid, p, init=None):
    inf = ctx.ligand_info(lid)
    n, T, H = p.chains, inf.n_torsions, inf.n_heavy
    ptrs = [ctx.alloc(n * 4), ctx.alloc(n * (7 + T) * 4), ctx.alloc(n * H * 12)]
    d_init = None
    try:
        if init is not None:
            d_init = ctx.alloc(init.nbytes)
            ctx.lib.vd_copy(ctx.h, d_init, np.ascontiguousarray(init, dtype=np.float32).ctypes.data, init.nbytes, 0)
        st = ctx.dock_noalloc(lid, p, *ptrs, d_init=d_init)
        out = (ctx.to_host(ptrs[0], (n,)), ctx.to_host(ptrs[1], (n, 7 + T)), ctx.to_host(ptrs[2], (n, H, 3)))
    finally:
        for q in ptrs + ([d_init] if d_init else []):
            ctx.free(q)
    return out, st


def test_noalloc_matches_dock(ctx):
    import vdock
    from vdock import workloads as W
    say("no-alloc: small scene")
    W.small_scene(vdock, ctx)
    big = ctx.add_ligand(vdock.synth_ligand(40, 10, 71))
    small = ctx.add_ligand(vdock.synth_ligand(14, 3, 72))
    cases = [(big, vdock.default_params(chains=512, search_depth=3, bfgs_steps=8, seed=4)),
             (small, vdock.default_params(chains=96, search_depth=3, bfgs_steps=8, seed=5)),
             (small, vdock.default_params(chains=64, search_depth=40, bfgs_steps=8, seed=6, policy=vdock.VD_MC_QVINA2))]
    for lid, p in cases:
        say("no-alloc: ligand", lid, "chains", p.chains, "policy", p.policy)
        (e, c, h), st = _noalloc(ctx, lid, p)
        e2, c2, h2, st2 = ctx.dock(lid, p)
        assert np.isfinite(e).all() and st.evals == st2.evals and st.evals > p.chains
        assert np.array_equal(e, e2) and np.array_equal(c, c2) and np.array_equal(h, h2)
    # given starting conformations
    inf = ctx.ligand_info(small)
    rng = np.random.default_rng(3)
    init = np.zeros((32, 7 + inf.n_torsions), dtype=np.float32)
    init[:, :3] = rng.uniform(-3, 3, (32, 3))
    q = rng.normal(size=(32, 4))
    init[:, 3:7] = q / np.linalg.norm(q, axis=1, keepdims=True)
    init[:, 7:] = rng.uniform(-np.pi, np.pi, (32, inf.n_torsions))
    p = vdock.default_params(chains=32, search_depth=2, bfgs_steps=8, seed=7)
    (e, c, h), _ = _noalloc(ctx, small, p, init=init)
    e2, c2, h2, _ = ctx.dock(small, p, init_conf=init)
    assert np.array_equal(e, e2) and np.array_equal(c, c2) and np.array_equal(h, h2)
    # beyond the scratch array: QuickVina2 histories of 4096 chains of a 10-torsion ligand
    p = vdock.default_params(chains=4096, search_depth=2, bfgs_steps=8, seed=8, policy=vdock.VD_MC_QVINA2)
    with pytest.raises(vdock.VdockError) as ei:
        _noalloc(ctx, big, p)
    assert ei.value.code == vdock.VD_ERR_CAPACITY
