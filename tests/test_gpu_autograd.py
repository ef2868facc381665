"""Forces and stress as adjoints (qbg.autograd over the device path) against
central finite differences of the device total energy at fixed coefficients,
on two small cells and on the 128-atom orthorhombic 28 x 30 x 32 bohr
workload (ecut 60 Ry, 256 bands):
position steps of 1e-4 bohr and strain steps of 1e-4, with a reference cell
fixing the basis (test/si2stress/check_stress.sh:20-90 for the strain
convention)."""
import numpy as np
import pytest
import torch

from qbg import configs as cf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import qbg
    c = qbg.Context(0)
    yield c
    c.close()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("idx", [11, 12, 4])
def test_autograd_forces_stress_vs_fd(ctx, idx):
    import qbg
    from qbg import autograd as qa
    s = cf.config(idx)
    # the 128-atom orthorhombic workload (144 x 160 x 168 grid) and two small cells
    ref = s.cell * (1.0 if idx == 4 else 1.04)
    ef0 = qbg.EnergyFunctional(ctx, s, refcell=ref)
    c = cf.gram_numpy(cf.random_coefficients(321 + idx, ef0.basis(0)["g2"], s.nst))
    ef0.release()
    model = qa.QboxModel(ctx, s, c, refcell=ref)
    tau = torch.tensor(s.tau, dtype=torch.float64, requires_grad=True)
    eps = torch.zeros((3, 3), dtype=torch.float64, requires_grad=True)
    e = qa.energy(tau, eps, model)
    e.backward()
    t0 = s.tau.copy()
    z = np.zeros((3, 3))
    h = 1e-4
    print(f"[case] config {idx}: E = {e.item():.12f}", flush=True)
    for (a, j) in [(0, 0), (0, 2), (s.nat - 1, 1)]:
        tp, tm = t0.copy(), t0.copy()
        tp[a, j] += h
        tm[a, j] -= h
        fd = (model.evaluate(tp, z)[0] - model.evaluate(tm, z)[0]) / (2 * h)
        g = tau.grad[a, j].item()
        print(f"  dE/dtau[{a},{j}] adjoint {g: .10e} fd {fd: .10e}")
        assert abs(g - fd) <= 1e-6 * max(abs(fd), 1e-3)
    for (i, j) in [(0, 0), (1, 2), (0, 1)]:
        u = np.zeros((3, 3))
        u[i, j] = u[j, i] = h
        fd = (model.evaluate(t0, u)[0] - model.evaluate(t0, -u)[0]) / (2 * h)
        g = eps.grad[i, j].item() + (eps.grad[j, i].item() if i != j else 0.0)
        print(f"  dE/deps[{i},{j}] adjoint {g: .10e} fd {fd: .10e}")
        assert abs(g - fd) <= 1e-6 * max(abs(fd), 1e-2)
