"""torch.autograd binding of the parameter adjoint (mspms_energy_param_grad).

`NonbondedEnergy.apply(pos, epsilon, sigma, charge, engine)` returns the
nonbonded energy uinter of every replica of `engine` (FP64, reduced units)
as a function of the packed positions (3, N) and of the per-site template
parameters (arrays over global site ids, reduced units).  Its backward pass
gives dE/dpos = -F and dE/d(epsilon, sigma, charge) from the HIP adjoint
kernels, so that force-field parameters can be fitted with torch optimisers.
Parameters of an atom type shared by several sites are expressed by indexing
a per-type tensor (autograd then sums the site derivatives).

The evaluation runs on the GPU through the C ABI; the tensors crossing this
function live on the host (the adjoint entry point takes host buffers).
"""
from __future__ import annotations

import numpy as np
import torch


class NonbondedEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, epsilon, sigma, charge, engine):
        engine.set_site_params(epsilon.detach().cpu().double().numpy(),
                               sigma.detach().cpu().double().numpy(),
                               charge.detach().cpu().double().numpy())
        p = pos.detach().cpu().double().numpy()
        f, terms, g = engine.param_grad(p[0], p[1], p[2])
        terms = terms if isinstance(terms, list) else [terms]
        energy = torch.tensor([t["uinter"] for t in terms], dtype=torch.float64)
        ctx.save_for_backward(torch.from_numpy(f), torch.from_numpy(g["epsilon"]),
                              torch.from_numpy(g["sigma"]), torch.from_numpy(g["charge"]))
        ctx.rep_atom0 = getattr(engine, "rep_atom0", np.array([0, p.shape[1]]))
        ctx.devices = (pos.device, epsilon.device, sigma.device, charge.device)
        return energy.to(pos.device)

    @staticmethod
    def backward(ctx, grad_e):
        f, ge, gs, gq = ctx.saved_tensors
        w = grad_e.detach().cpu().double()
        # dE_r/dpos = -F on the atoms of replica r
        a = ctx.rep_atom0
        scale = torch.repeat_interleave(w, torch.as_tensor(np.diff(a)))
        dpos = -f * scale[None, :]
        d = ctx.devices
        return (dpos.to(d[0]), (w[:, None] * ge).sum(0).to(d[1]),
                (w[:, None] * gs).sum(0).to(d[2]), (w[:, None] * gq).sum(0).to(d[3]), None)


def site_params(system):
    """The template epsilon, sigma, charge of `system` as three float64
    tensors over global site ids."""
    cat = lambda name: torch.tensor(np.concatenate([getattr(s, name) for s in system.species]),
                                    dtype=torch.float64)
    return cat("epsilon"), cat("sigma"), cat("charge")
