"""torch.autograd.Function over the C ABI: the total energy of a Gamma-point
system as a function of the atomic positions and of a homogeneous strain, at
fixed plane-wave coefficients, with the library's force and stress kernels as
its adjoints.

    E(tau, eps):  positions tau (nat x 3, bohr); the cell a_i -> (1 + eps) a_i
                  and tau -> (1 + eps) tau (fractional coordinates fixed);
                  eps is a symmetric 3 x 3 strain.
    dE/dtau       = -F          (EnergyFunctional::energy forces,
                                 src/EnergyFunctional.cpp:841-930 and
                                 src/NonLocalPotential.cpp:1634-1784)
    dE/deps_ij    = -omega sigma_ij   (the stress tensor, sigma in
                                 src/EnergyFunctional.cpp:932-1041; for the
                                 symmetric eps the off-diagonal derivative is
                                 split evenly between ij and ji)

The basis is fixed by a reference cell (Wavefunction's refcell,
src/Basis.cpp:314-325), so the strained systems share the coefficients index
by index.  Gradients are those of the unstrained state: the forward pass
evaluates at the given eps, the backward pass returns -F and -omega sigma of
that state, which is the derivative with respect to a further strain applied
to the strained cell (exact at eps = 0, the usual use).
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from ._lib import Context, DArray, EnergyFunctional

VOIGT = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]


def sigma_matrix(v):
    """Voigt (xx yy zz xy yz xz) -> symmetric 3 x 3."""
    s = np.zeros((3, 3))
    for k, (i, j) in enumerate(VOIGT):
        s[i, j] = s[j, i] = v[k]
    return s


class QboxModel:
    """One system on the device: coefficients resident, one energy functional
    moved to each new cell (qbg_ef_cell_moved over the reference cell) and to
    each new set of positions (atoms_moved)."""

    def __init__(self, ctx: Context, sysdef, coefficients: np.ndarray, refcell=None,
                 prec: str = "fp64"):
        self.ctx = ctx
        self.sysdef = sysdef
        self.refcell = np.asarray(refcell if refcell is not None else sysdef.cell, dtype=float)
        self.prec = prec
        self.c = ctx.to_device(np.ascontiguousarray(coefficients, dtype=np.complex128))
        self._ef = None
        self._cell = None

    def _functional(self, cell, tau):
        if self._ef is None:
            sd = dataclasses.replace(self.sysdef, cell=cell, tau=tau)
            self._ef = EnergyFunctional(self.ctx, sd, refcell=self.refcell)
            self._cell = cell.copy()
            return self._ef
        if not np.array_equal(cell, self._cell):
            self._ef.cell_moved(cell)
            self._cell = cell.copy()
        self._ef.set_positions(tau)
        return self._ef

    def evaluate(self, tau: np.ndarray, eps: np.ndarray):
        """(etotal, fion (nat x 3), sigma (3 x 3), omega) at positions tau and
        strain eps (applied to the cell and to tau)."""
        F = np.eye(3) + eps
        cell = np.asarray(self.sysdef.cell, dtype=float) @ F.T
        tau_s = np.asarray(tau, dtype=float) @ F.T
        ef = self._functional(cell, tau_s)
        r = ef.evaluate(self.c, None, forces=True, stress=True, prec=self.prec)
        omega = abs(np.linalg.det(cell))
        return r["etotal"], np.asarray(r["fion"]), sigma_matrix(r["sigma"][0]), omega


class EnergyFn(torch.autograd.Function):
    """E = EnergyFn.apply(tau, eps, model): tau (nat, 3) and eps (3, 3) are
    float64 tensors; returns a 0-d float64 tensor."""

    @staticmethod
    def forward(ctx, tau, eps, model: QboxModel):
        e, f, sig, omega = model.evaluate(tau.detach().cpu().numpy(),
                                          eps.detach().cpu().numpy())
        ctx.save_for_backward(torch.as_tensor(f, dtype=torch.float64),
                              torch.as_tensor(sig, dtype=torch.float64))
        ctx.omega = omega
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        f, sig = ctx.saved_tensors
        return -grad * f, -grad * ctx.omega * sig, None


def energy(tau: torch.Tensor, eps: torch.Tensor, model: QboxModel) -> torch.Tensor:
    return EnergyFn.apply(tau, eps, model)
