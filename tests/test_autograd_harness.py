"""The torch.autograd.Function harness (qbg.autograd) on the CPU with a model
whose energy, forces and stress are known in closed form: the gradients that
backward returns must match central finite differences of forward."""
import numpy as np
import pytest
import torch

from qbg import autograd as qa


class QuadraticModel:
    """E = 1/2 sum k (tau - tau0)^2 + omega0 det(1 + eps) p; forces -k (tau - tau0),
    stress sigma = -(1/omega) dE/deps = -p I for eps = 0 (omega = omega0)."""

    def __init__(self):
        rng = np.random.default_rng(3)
        self.tau0 = rng.standard_normal((4, 3))
        self.k = 0.7
        self.p = 0.013
        self.omega0 = 123.0

    def evaluate(self, tau, eps):
        F = np.eye(3) + eps
        omega = self.omega0 * np.linalg.det(F)
        d = tau - self.tau0
        e = 0.5 * self.k * np.sum(d * d) + self.p * omega
        # -(1/omega) d/deps' of p omega(1 + eps') det = -p I at the current cell
        return e, -self.k * d, -self.p * np.eye(3), omega


def test_backward_matches_finite_differences():
    m = QuadraticModel()
    tau = torch.tensor(m.tau0 + 0.1, dtype=torch.float64, requires_grad=True)
    eps = torch.zeros((3, 3), dtype=torch.float64, requires_grad=True)
    e = qa.energy(tau, eps, m)
    e.backward()
    h = 1e-5
    t0 = tau.detach().numpy()
    for (i, j) in [(0, 0), (2, 1), (3, 2)]:
        tp, tm = t0.copy(), t0.copy()
        tp[i, j] += h
        tm[i, j] -= h
        fd = (m.evaluate(tp, np.zeros((3, 3)))[0] - m.evaluate(tm, np.zeros((3, 3)))[0]) / (2 * h)
        assert abs(tau.grad[i, j].item() - fd) < 1e-8
    for (i, j) in [(0, 0), (1, 1), (0, 1)]:
        u = np.zeros((3, 3))
        u[i, j] = u[j, i] = h
        fd = (m.evaluate(t0, u)[0] - m.evaluate(t0, -u)[0]) / (2 * h)
        g = eps.grad[i, j].item() + (eps.grad[j, i].item() if i != j else 0.0)
        assert abs(g - fd) < 1e-8


def test_sigma_matrix_voigt_order():
    s = qa.sigma_matrix([1, 2, 3, 4, 5, 6])
    assert np.array_equal(s, [[1, 4, 6], [4, 2, 5], [6, 5, 3]])
