"""torch.autograd over the HIP adjoint of the nonbonded path.

`nonbonded_energy(engine, charge, lj_A, lj_B)` returns the total nonbonded
energy (LJ + long-range dispersion + PME direct + exclusion correction +
reciprocal + self) of the engine's current coordinates, differentiable with
respect to the per-atom charges and the stored (x12 / x6) LJ pair tables.
`force_matching_loss(engine, charge, lj_A, lj_B, target)` returns
L = sum_i |F_i - target_i|^2 with the same parameters. Backward calls the
adjoint kernels (sponge_nb_energy_gradient / sponge_nb_force_matching_gradient),
which recompute the pair terms instead of storing them.

Charges are FP32 device tensors in SPONGE's units (e x 18.2223); tables are
host-side float arrays in the layout of sponge_lj_create. Gradients come back
in FP64 and are cast to the inputs' dtypes. `project_net_charge` removes the
component that would change the total charge.
"""
from __future__ import annotations

import numpy as np
import torch

from . import nb


def _set_params(engine, charge, lj_A, lj_B):
    engine.charge.copy_(charge.detach().to(engine.charge.dtype))
    A = lj_A.detach().cpu().numpy().astype(np.float32)
    B = lj_B.detach().cpu().numpy().astype(np.float32)
    engine.lj.set_tables(A, B)


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, charge, lj_A, lj_B, engine):
        _set_params(engine, charge, lj_A, lj_B)
        engine.crd_to_uint()
        engine.force_step(need_energy=1, need_virial=0)
        p = engine.energy_parts()
        e = p["lj"] + p["lrc"] + p["direct"] + p["correction"] + p["self"] + p["reciprocal"]
        ctx.engine = engine
        ctx.dtypes = (charge.dtype, lj_A.dtype, lj_B.dtype, lj_A.device)
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        eng = ctx.engine
        P = len(eng.sys.lj_A)
        gq = torch.zeros(eng.n, dtype=torch.float64, device=eng.device)
        gA = torch.zeros(P, dtype=torch.float64, device=eng.device)
        gB = torch.zeros(P, dtype=torch.float64, device=eng.device)
        eng.energy_gradient(gq, gA, gB)
        torch.cuda.synchronize()
        g = g.to(torch.float64).item()
        qd, ad, bd, dev = ctx.dtypes
        return (gq * g).to(qd), (gA * g).to(device=dev, dtype=ad), (gB * g).to(device=dev, dtype=bd), None


class _ForceLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, charge, lj_A, lj_B, target, engine):
        _set_params(engine, charge, lj_A, lj_B)
        engine.force_step(0, 0, from_crd=True, overwrite=True)
        frc = engine.frc.clone()
        tgt = target.detach().to(torch.float32).contiguous()
        P = len(engine.sys.lj_A)
        loss = torch.zeros(1, dtype=torch.float64, device=engine.device)
        gq = torch.zeros(engine.n, dtype=torch.float64, device=engine.device)
        gA = torch.zeros(P, dtype=torch.float64, device=engine.device)
        gB = torch.zeros(P, dtype=torch.float64, device=engine.device)
        ctx.engine = engine
        ctx.save_for_backward(frc, tgt)
        ctx.dtypes = (charge.dtype, lj_A.dtype, lj_B.dtype, lj_A.device, target.dtype)
        ctx.buffers = (loss, gq, gA, gB)
        return ((frc.to(torch.float64) - tgt.to(torch.float64)) ** 2).sum()

    @staticmethod
    def backward(ctx, g):
        eng = ctx.engine
        frc, tgt = ctx.saved_tensors
        loss, gq, gA, gB = ctx.buffers
        eng.force_matching_gradient(frc, tgt, loss, gq, gA, gB)
        g = g.to(torch.float64)
        qd, ad, bd, dev, td = ctx.dtypes
        gT = (-2.0 * (frc.to(torch.float64) - tgt.to(torch.float64)) * g).to(td)
        return ((gq * g).to(qd), (gA * g).to(device=dev, dtype=ad), (gB * g).to(device=dev, dtype=bd), gT,
                None)


def nonbonded_energy(engine: "nb.NonbondedEngine", charge, lj_A, lj_B):
    return _Energy.apply(charge, lj_A, lj_B, engine)


def force_matching_loss(engine: "nb.NonbondedEngine", charge, lj_A, lj_B, target):
    return _ForceLoss.apply(charge, lj_A, lj_B, target, engine)


def project_net_charge(grad_charge):
    """Gradient restricted to charge changes of zero total (sum dq = 0)."""
    return grad_charge - grad_charge.mean()
