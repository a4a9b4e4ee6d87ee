"""torch.autograd wrapper of the Davidson energies and their adjoint.

``EnergyFunction.apply(valone, valtwo, valso, ham, nroots)`` returns the k
lowest eigenvalues of H(valone, valtwo) (GPU block Davidson through the C
ABI); its backward calls ``psoci_rdm`` with the incoming gradient as the root
weights, returning dL/dvalone, dL/dvaltwo (the generalised 1- and 2-RDM
contracted with dL/dE) and dL/dvalso (zero: with one configuration symmetry
the reference never enters its spin-orbit branch,
PsociHamiltonian.C:223-238).  Tensors are CPU float64; the compute runs on
the GPU the Hamiltonian context lives on.
"""
from __future__ import annotations

import numpy as np
import torch

from . import api
from .inputs import Integrals


class EnergyFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, valone, valtwo, valso, ham, nroots, mode=api.DIRECT, tol=1e-8):
        ints = Integrals(nbf=ham.cfg.nbf, valone=valone.detach().cpu().numpy().astype(np.float64),
                         valtwo=valtwo.detach().cpu().numpy().astype(np.float64))
        ham.set_integrals(ints)
        if mode == api.STORED:
            ham.build()
        out = ham.davidson(nroots, prec=api.FP64, mode=mode, tol=tol)
        if not out["converged"]:
            raise RuntimeError("Davidson did not converge")
        ctx.ham = ham
        ctx.evecs = out["evecs"]
        ctx.nso = valso.numel()
        return torch.from_numpy(out["evals"].copy())

    @staticmethod
    def backward(ctx, grad_e):
        w = grad_e.detach().cpu().numpy().astype(np.float64)
        g1, g2, gso = ctx.ham.rdm(ctx.evecs, w)
        return (torch.from_numpy(g1), torch.from_numpy(g2), torch.from_numpy(gso[:ctx.nso].copy()), None, None,
                None, None)


def energies(valone, valtwo, valso, ham, nroots, mode=api.DIRECT, tol=1e-8):
    return EnergyFunction.apply(valone, valtwo, valso, ham, nroots, mode, tol)
