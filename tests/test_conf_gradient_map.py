"""The map from the adjoint's change coordinates to a conformation's own coordinates, used by the
torch.autograd wrapper of the GPU tests (CPU only, on the oracle).

The gradient of eval_deriv is in the change coordinates the optimiser uses (V11/lib/conf.h:
60-66, 118-124): force on the root, torque about the root origin, one derivative per torsion.
For a conformation (position, unit quaternion q = (w, v), torsions) the position and torsion
parts are derivatives already; a small rotation d moves the quaternion by (0, d/2) (x) q, so
dE/dq = 2 (0, torque) (x) q, tangent to the unit sphere.  Pinned by FP64 central differences of
the oracle's energy on the reference's 1u1b scene (h = 1e-6, the quaternion perturbed and then
renormalised, as the kinematics normalises it), with the intermolecular term alone: the
intramolecular gradient uses the tabulated dor, which is not the exact derivative of the
interpolated energy (its own check compares it with the oracle's dor-based gradient)."""
import os

import numpy as np

import oracle as orc
from conftest import GOLDEN, golden

CENTER = [-30.661, -29.540, 17.010]   # V11/input_file_example/1u1b_config.txt
SIZE = [24.0, 26.25, 22.5]


def quat_grad(q, torque):
    """dE/dq = 2 (0, torque) (x) q (Hamilton product) for q = (w, x, y, z)."""
    w, v = q[..., :1], q[..., 1:]
    return 2 * np.concatenate([-(torque * v).sum(-1, keepdims=True), w * torque + np.cross(torque, v)], -1)


def conf_grad(conf, g):
    """The change-coordinate gradient g (6 + T) in the conformation's coordinates (7 + T)."""
    return np.concatenate([g[..., :3], quat_grad(conf[..., 3:7], g[..., 3:6]), g[..., 6:]], -1)


def test_conf_gradient_map_matches_fp64_central_differences():
    xyz, xs = orc.parse_receptor(golden("1u1b_protein.pdbqt"))[:2]
    b, e, n = orc.grid_dims(CENTER, SIZE)
    g = orc.Grid(orc.Lut(), b, e, n)
    g.set_receptor(xyz, xs)
    g.populate(True)
    lig = orc.Ligand(golden("1u1b_ligand.pdbqt"))
    fx = np.load(os.path.join(GOLDEN, "oracle_eval_1u1b.npz"), allow_pickle=False)
    h = 1e-6
    orc.set_terms(inter=True, intra=False)
    try:
        check_modes(g, lig, fx["conf"][:6], h)
    finally:
        orc.set_terms(True, True)


def check_modes(g, lig, confs, h):
    for c in confs:                           # the reference's docked modes
        _, gr, _, _ = orc.eval_deriv(g, lig, c)
        an = conf_grad(c, gr)
        fd = np.zeros_like(c)
        for j in range(len(c)):
            cp, cm = c.copy(), c.copy()
            cp[j] += h
            cm[j] -= h
            for x in (cp, cm):
                x[3:7] /= np.linalg.norm(x[3:7])
            fd[j] = (orc.eval_deriv(g, lig, cp)[0] - orc.eval_deriv(g, lig, cm)[0]) / (2 * h)
        tol = 1e-4 * max(1.0, np.abs(an).max()) + 1e-4
        assert np.all(np.abs(fd - an) <= tol), (np.abs(fd - an).max(), tol)
        # a slip in the map (the factor 2, the order of the product) misses by far more
        wrong = np.concatenate([gr[:3], quat_grad(c[3:7], gr[3:6]) / 2, gr[6:]])
        assert np.abs(fd - wrong).max() > 10 * tol
