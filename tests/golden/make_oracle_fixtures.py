"""Writes the oracle fixtures of tests/golden/ (run from the repository root):

  oracle_lut_slices.npz   smooth value, smooth dor and fast value of four type pairs, all bins
  oracle_maps_1u1b.npz    the 17 maps of a 9^3 patch (3 A box at 0.375 A) at the centre of the
                          1u1b search box (V11/input_file_example/1u1b_config.txt)
  oracle_eval_1u1b.npz    64 conformations of the 1u1b ligand in the example's search box:
                          the reference's nine docked modes (V11/docking_result/
                          1u1b_docking_result.pdbqt, fitted to conformations), 39 seeded
                          perturbations of them and 16 seeded random poses; energy and gradient
                          (v = 10, slope 1e6)

All from the FP64 oracle (oracle/vina_oracle.cpp), so that a change of the oracle, or of the
device code against it, shows against frozen numbers.  Inputs are the reference's example files.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle as orc  # noqa: E402
from test_oracle_rescore import conf_from_coords, parse_models  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")
CENTER = [-30.661, -29.540, 17.010]        # V11/input_file_example/1u1b_config.txt
SIZE = [24.0, 26.25, 22.5]
PAIRS = [(0, 0), (2, 8), (5, 12), (1, 16)]   # C_H-C_H, N_D-O_A, N_DA-F_H, C_P-Met_D


def read(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def confs(lig, n, seed, center, spread):
    rng = np.random.default_rng(seed)
    c = np.zeros((n, 7 + lig.T))
    c[:, :3] = np.asarray(center) + rng.uniform(-spread, spread, (n, 3))
    q = rng.normal(size=(n, 4))
    c[:, 3:7] = q / np.linalg.norm(q, axis=1, keepdims=True)
    c[:, 7:] = rng.uniform(-np.pi, np.pi, (n, lig.T))
    return c


def fp32(c):
    return np.asarray(c).astype(np.float32).astype(np.float64)   # exactly representable in the FP32 ABI


def docked_confs(lig, seed):
    """The reference's modes, then perturbations of them (shift 0.5 A, a rotation of about
    0.1 rad, torsions 0.25 rad)."""
    models, _ = parse_models(read("1u1b_docking_result.pdbqt"))
    ser = lig.serials()
    modes = [conf_from_coords(lig, np.array([m[s] for s in ser])) for m in models]
    rng = np.random.default_rng(seed)
    out = list(modes)
    for k in range(39):
        c = modes[k % len(modes)].copy()
        c[:3] += rng.normal(0, 0.5, 3)
        dq = np.concatenate([[1.0], rng.normal(0, 0.05, 3)])
        w1, v1, w2, v2 = dq[0], dq[1:], c[3], c[4:7]
        q = np.concatenate([[w1 * w2 - v1 @ v2], w1 * v2 + w2 * v1 + np.cross(v1, v2)])
        c[3:7] = q / np.linalg.norm(q)
        c[7:] += rng.normal(0, 0.25, lig.T)
        out.append(c)
    return out


def build():
    lut = orc.Lut()
    tp = [orc.tri_index(a, b) for a, b in PAIRS]
    se, sd, fa = zip(*(lut.tables(t) for t in tp))
    out = {"lut": dict(pairs=np.array(PAIRS, dtype=np.int32), tp=np.array(tp, dtype=np.int32),
                       smooth_e=np.array(se), smooth_dor=np.array(sd), fast=np.array(fa))}
    xyz, xs = orc.parse_receptor(read("1u1b_protein.pdbqt"))[:2]
    b, e, n = orc.grid_dims(CENTER, [3.0, 3.0, 3.0])
    g = orc.Grid(lut, b, e, n)
    g.set_receptor(xyz, xs)
    g.populate(True)
    out["maps"] = dict(begin=np.array(b), end=np.array(e), n=np.array(n, dtype=np.int32), maps=g.maps())
    b, e, n = orc.grid_dims(CENTER, SIZE)
    gb = orc.Grid(lut, b, e, n)
    gb.set_receptor(xyz, xs)
    gb.populate(True)
    lig = orc.Ligand(read("1u1b_ligand.pdbqt"))
    c = fp32(docked_confs(lig, 2024) + list(confs(lig, 16, 2024, CENTER, 2.0)))
    es, gs = [], []
    for k in range(len(c)):
        en, gr, _, _ = orc.eval_deriv(gb, lig, c[k])
        es.append(en)
        gs.append(gr)
    out["eval"] = dict(begin=np.array(b), end=np.array(e), n=np.array(n, dtype=np.int32), conf=c,
                       e=np.array(es), grad=np.array(gs))
    return out


if __name__ == "__main__":
    out = build()
    np.savez_compressed(os.path.join(GOLDEN, "oracle_lut_slices.npz"), **out["lut"])
    np.savez_compressed(os.path.join(GOLDEN, "oracle_maps_1u1b.npz"), **out["maps"])
    np.savez_compressed(os.path.join(GOLDEN, "oracle_eval_1u1b.npz"), **out["eval"])
    print("written")
