"""torch.autograd wrapper over the adjoint kernel (cass_adjoint).

BoxEnergy(eps, sigma, charges, positions) returns the nonbonded energy of a
box (intermolecular LJ + real-space q-q, intramolecular scaled terms and the
periodic image part, Ewald reciprocal and self energies, LJ tail) as a
float64 tensor; its backward uses the gradients the HIP adjoint kernel
returns in the same launch: dE/d eps_t and dE/d sigma_t per atom type (through
Lorentz-Berthelot mixing), dE/d q per species-atom kind, and -forces for the
positions.  Units are the engine's: eps in amu A^2/ps^2 (K * 0.8314472),
sigma in A, q in e, energy in amu A^2/ps^2.

Typical use is reweighting-based force-field fitting: frames are uploaded
once per evaluation, parameters carry requires_grad.
"""
from __future__ import annotations

import numpy as np
import torch

from .systems import TINY_NUMBER


def lorentz_berthelot(eps_t: np.ndarray, sig_t: np.ndarray):
    """vdw_param1/2_table from per-type parameters, as create_nonbond_table.f90:370-392."""
    nt = len(eps_t)
    eps = np.zeros((nt, nt))
    sig = np.zeros((nt, nt))
    for i in range(nt):
        for j in range(nt):
            if i == j:
                eps[i, j], sig[i, j] = eps_t[i], sig_t[i]
                continue
            ei, ej = eps_t[i], eps_t[j]
            eps[i, j] = 0.0 if (ei <= TINY_NUMBER or ej <= TINY_NUMBER) else np.sqrt(ei * ej)
            si, sj = sig_t[i], sig_t[j]
            sig[i, j] = 0.0 if (si <= TINY_NUMBER or sj <= TINY_NUMBER) else (si + sj) * 0.5
    return eps, sig


class BoxEnergy(torch.autograd.Function):
    """E = BoxEnergy.apply(eps_t, sig_t, q_kind, positions, box)."""

    @staticmethod
    def forward(ctx, eps_t, sig_t, q_kind, positions, box):
        eps_tab, sig_tab = lorentz_berthelot(eps_t.detach().cpu().numpy().astype(np.float64),
                                             sig_t.detach().cpu().numpy().astype(np.float64))
        eng = box.eng
        eng.set_nonbond_params(eps_tab, sig_tab, q_kind.detach().cpu().numpy().astype(np.float64))
        if positions is not None:
            box.upload(box.spec, box.live, positions.detach().cpu().numpy().astype(np.float64).ravel())
        if box.system.charge_style == "ewald":
            box.ewald_setup()
        e, de, ds, dq, f = box.adjoint()
        dev = eps_t.device
        ctx.has_positions = positions is not None
        ctx.save_for_backward(torch.as_tensor(de, device=dev), torch.as_tensor(ds, device=dev),
                              torch.as_tensor(dq, device=dev), torch.as_tensor(-f, device=dev))
        return torch.tensor(e, dtype=torch.float64, device=dev)

    @staticmethod
    def backward(ctx, g):
        de, ds, dq, dr = ctx.saved_tensors
        return g * de, g * ds, g * dq, (g * dr if ctx.has_positions else None), None


def box_energy(box, eps_t, sig_t, q_kind, positions=None):
    return BoxEnergy.apply(eps_t, sig_t, q_kind, positions, box)
