"""Synthetic workloads of the docking benchmark, built through libvdock's own generators
(vd_synth_receptor, vd_synth_ligand): one function per configuration, shared by bench.py and the
parity tests so that both run the same inputs.

  single_ligand_scene   8,000-atom receptor (60 A cube, 10 A cavity), 30 A box at 0.375 A (81^3)
  small_scene           2,000-atom receptor (36 A cube, 8 A cavity), 20 A box (55^3)
  large_box_scene       20,000-atom receptor (90 A cube, 25 A cavity), 70 A box (188^3)
  screening_specs       1,024 ligands: heavy atoms uniform in [10, 60], torsions uniform in
                        [0, min(16, heavy / 3)], ligand i generated with seed i
  large_box_specs       4,096 ligands: heavy atoms uniform in [10, 80], torsions uniform in [0, 24]
"""
import numpy as np


def _scene(vdock, ctx, n_atoms, cube, cavity, seed, box_size):
    xyz, xs = vdock.synth_receptor(n_atoms, cube, cavity, seed)
    box = vdock.box_from_center([0, 0, 0], [box_size] * 3, 0.375)
    if ctx is not None:
        ctx.set_receptor(xyz, xs)
        ctx.grid_populate(box)
    return xyz, xs, box


def small_scene(vdock, ctx=None):
    """2,000 heavy atoms, 36 A cube, 8 A cavity, seed 42; 20 A box -> 55^3 points."""
    return _scene(vdock, ctx, 2000, 36.0, 8.0, 42, 20.0)


def single_ligand_scene(vdock, ctx=None):
    """8,000 heavy atoms, 60 A cube, 10 A cavity, seed 42; 30 A box -> 81^3 points."""
    return _scene(vdock, ctx, 8000, 60.0, 10.0, 42, 30.0)


def large_box_scene(vdock, ctx=None):
    """20,000 heavy atoms, 90 A cube, 25 A cavity, seed 4242; 70 A box -> 188^3 points."""
    return _scene(vdock, ctx, 20000, 90.0, 25.0, 4242, 70.0)


def screening_specs(n=1024):
    """(heavy atoms, torsions, seed) of the batched-screening ligands."""
    rng = np.random.default_rng(2024)
    out = []
    for i in range(n):
        heavy = int(rng.integers(10, 61))
        tors = int(rng.integers(0, min(16, heavy // 3) + 1))
        out.append((heavy, tors, i))
    return out


def large_box_specs(n=4096):
    """(heavy atoms, torsions, seed) of the large-box screening ligands (the torsions drawn are
    capped by the rotatable bonds the generated tree has)."""
    rng = np.random.default_rng(4096)
    out = []
    for i in range(n):
        heavy = int(rng.integers(10, 81))
        tors = int(rng.integers(0, 25))
        out.append((heavy, tors, 100000 + i))
    return out


def texts(vdock, specs):
    return [vdock.synth_ligand(h, t, s) for h, t, s in specs]
