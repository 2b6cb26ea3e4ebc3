"""The hand-written forward/adjoint pair through torch.autograd: a thin torch.autograd.Function
over the C ABI's evaluation (vd_eval; the forward and the adjoint run on the GPU inside the
library), whose backward returns the device's adjoint gradient in the conformation's own
coordinates (position, quaternion, torsions; the map from the adjoint's force / torque /
torsion coordinates and its FP64 central-difference check are in test_conf_gradient_map.py).
Checked on the reference's 1u1b scene against the FP64 oracle's gradient under the same map.

The tensors stay on the host: the torch wheel carries its own HIP and HSA runtimes, and a second
HSA runtime in a process that already uses the GPU through libvdock's finds no device, so torch's
GPU side is not initialised here."""
import os

import numpy as np
import pytest

import oracle as orc
from conftest import GOLDEN, golden
from test_conf_gradient_map import CENTER, SIZE, conf_grad

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_function():
    import torch

    class VinaEnergy(torch.autograd.Function):
        """E(conf) for a batch of conformations [n, 7 + T] (float32) of ligand lid, computed by the
        library on the GPU; dE/dq = 2 (0, torque) (x) q for the quaternion part."""

        @staticmethod
        def forward(fctx, conf, vd, lid):
            c = conf.detach().contiguous().to(torch.float32)
            e, g = vd.eval(lid, c.numpy())
            g = torch.from_numpy(g)
            fctx.save_for_backward(c, g)
            return torch.from_numpy(e)

        @staticmethod
        def backward(fctx, grad_e):
            conf, g = fctx.saved_tensors
            q, tq = conf[:, 3:7], g[:, 3:6]
            w, v = q[:, :1], q[:, 1:]
            gq = 2 * torch.cat([-(tq * v).sum(1, keepdim=True), w * tq + torch.cross(tq, v, dim=1)], 1)
            return torch.cat([g[:, :3], gq, g[:, 6:]], 1) * grad_e[:, None], None, None

    return VinaEnergy


def test_autograd_backward_matches_oracle(ctx):
    """autograd through the C ABI on 24 conformations near the reference's docked modes, weighted
    so that the backward's scaling by the incoming gradient shows: energies within
    max(1e-4, 1e-3 |e|) and conformation gradients within 1e-3 max(1, max|g|) + 1e-3 of the
    oracle's (FP32 device against FP64 on the same maps)."""
    import torch
    import vdock
    ctx.set_receptor_pdbqt(golden("1u1b_protein.pdbqt"))
    box = vdock.box_from_center(CENTER, SIZE, 0.375)
    ctx.grid_populate(box)
    g = orc.Grid(orc.Lut(), *box.as_tuple())
    g.set_maps(ctx.grid_download().astype(np.float64))       # the oracle evaluates on the device's maps
    lig = orc.Ligand(golden("1u1b_ligand.pdbqt"))
    lid = ctx.add_ligand(golden("1u1b_ligand.pdbqt"))
    fx = np.load(os.path.join(GOLDEN, "oracle_eval_1u1b.npz"), allow_pickle=False)
    lo, hi = np.asarray(box.as_tuple()[0]), np.asarray(box.as_tuple()[1])
    keep = [k for k in range(len(fx["conf"]))
            if fx["e"][k] < 100 and ((x := lig.set_conf(fx["conf"][k])) > lo).all() and (x < hi).all()]
    VinaEnergy = make_function()
    conf = torch.tensor(fx["conf"][keep[:24]], dtype=torch.float32, requires_grad=True)
    say("autograd over", len(conf), "conformations")
    e = VinaEnergy.apply(conf, ctx, lid)
    w = torch.linspace(0.5, 2.0, len(conf))
    (e * w).sum().backward()
    e_dev, g_dev, w_h = e.detach().numpy(), conf.grad.numpy(), w.numpy()
    worst = 0.0
    for i, c in enumerate(conf.detach().numpy().astype(np.float64)):
        e_ref, gr, _, _ = orc.eval_deriv(g, lig, c)
        an = conf_grad(c, gr) * w_h[i]
        assert abs(e_dev[i] - e_ref) <= max(1e-4, 1e-3 * abs(e_ref)), (i, e_dev[i], e_ref)
        tol = 1e-3 * max(1.0, np.abs(an).max()) + 1e-3
        assert np.all(np.abs(g_dev[i] - an) <= tol), (i, np.abs(g_dev[i] - an).max(), tol)
        worst = max(worst, np.abs(g_dev[i] - an).max() / max(1.0, np.abs(an).max()))
    say("autograd: worst relative gradient difference", worst)
