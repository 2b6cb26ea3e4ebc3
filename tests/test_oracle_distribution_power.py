"""Power of the chain-distribution criteria (helpers.compare_chain_distributions), shown on the CPU
with the oracle on both sides, on the CPU-runnable configuration (2,000-atom synthetic receptor,
55^3 box built by the oracle, a 20-atom, 5-torsion ligand at the reference's heuristic depth 5 and
16 BFGS steps, 32 seeds x 512 chains a side):

  - two independent sets of seeds pass (the criteria do not fail by chance);
  - each plausible mistake planted in one side fails them: the mutation amplitude 1 instead of 2
    (mutate.cpp:62-127), one BFGS step fewer than the heuristic (main.cpp:354-391), and the search
    curl cap 100 instead of hunt_cap's 10.

A Metropolis temperature of 1.0 and mutation over 1+T entities pass at this depth (five MC steps
decide little there); test_oracle_steps.py pins both step by step.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import compare_chain_distributions

SEEDS, CHAINS = 32, 512


@pytest.fixture(scope="module")
def scene():
    import vdock
    xyz, xs = vdock.synth_receptor(2000, 36.0, 8.0, 42)
    box = vdock.box_from_center([0, 0, 0], [20.0] * 3, 0.375)
    b, e, n = box.as_tuple()
    g = orc.Grid(orc.Lut(), b, e, n)
    g.set_receptor(xyz.astype(np.float64), xs)
    g.populate(True)
    lig = orc.Ligand(vdock.synth_ligand(20, 5, 1))
    depth, bfgs = vdock.heuristics(20, lig.T)
    assert (depth, bfgs) == (5, 16)
    return dict(grid=g, lig=lig, b=b, e=e, depth=depth, bfgs=bfgs)


def run(s, seeds, bfgs=None, **consts):
    G, gb, ev = [], [], 0
    for seed in seeds:
        p = orc.mc_params(CHAINS, s["depth"], bfgs or s["bfgs"], s["b"], s["e"], seed=seed, stream=0, threads=8)
        for k, v in consts.items():
            setattr(p, k, v)
        eo, _, _, n = orc.mc_run(s["grid"], s["lig"], p)
        G.append(np.asarray(eo))
        gb.append(float(eo.min()))
        ev += n
    return G, gb, ev


@pytest.fixture(scope="module")
def base(scene):
    return run(scene, range(SEEDS))


def test_independent_seeds_pass(scene, base):
    other = run(scene, range(100, 100 + SEEDS))
    r = compare_chain_distributions("independent seeds", base[0], other[0], base[1], other[1], base[2], other[2],
                                    check=False)
    assert r["ok"], r


@pytest.mark.parametrize("mistake", ["amplitude 1", "one BFGS step fewer", "curl cap 100"])
def test_planted_mistakes_fail(scene, base, mistake):
    kw = {"amplitude 1": dict(amplitude=1.0), "one BFGS step fewer": dict(bfgs=scene["bfgs"] - 1),
          "curl cap 100": dict(v_intra=100.0)}[mistake]
    other = run(scene, range(100, 100 + SEEDS), **kw)
    r = compare_chain_distributions(mistake, base[0], other[0], base[1], other[1], base[2], other[2], check=False)
    assert not r["ok"], r
