"""Host logic of the autograd Function (CPU, a stand-in for the GPU blob): backward
refuses a run that a later trajectory or material change replaced."""
import numpy as np
import pytest
import torch

from ffea_ci_demo_amd.adjoint import trajectory


class FakeBlob:
    """records the calls BlobTrajectory makes; positions = pos0 + sum of G"""

    def __init__(self, ne, nt):
        self.ne, self.nt = ne, nt
        self.run_generation = 0
        self.calls = []

    def set_material(self, A, B, G, E):
        self.run_generation += 1
        self.G = np.array(G)

    def set_positions(self, p):
        self.p0 = np.array(p)

    def run_checkpointed(self, nsteps, interval):
        self.run_generation += 1
        self.calls.append(("run", nsteps, float(self.G.sum())))

    def positions(self):
        return self.p0 + self.G.sum()

    def adjoint(self, seed):
        self.calls.append(("adjoint", float(self.G.sum())))
        z = np.zeros(self.ne)
        return z, z, np.full(self.ne, seed.sum()), z, np.zeros_like(seed)


def test_backward_of_replaced_run_refused():
    b = FakeBlob(4, 3)
    t = lambda v: torch.tensor(v, dtype=torch.float64, requires_grad=True)
    G, E, A, B = t(np.ones(4)), t(np.ones(4)), t(np.ones(4)), t(np.ones(4))
    x1 = trajectory(b, G, E, A, B, np.zeros((3, 3)), 5)
    x2 = trajectory(b, G * 2, E, A, B, np.zeros((3, 3)), 5)
    with pytest.raises(RuntimeError, match="another trajectory"):
        x1.sum().backward()
    assert not any(c[0] == "adjoint" for c in b.calls)
    x2.sum().backward()
    assert b.calls[-1] == ("adjoint", 8.0)
    assert np.allclose(G.grad.numpy(), 2 * 9.0)


def test_backward_in_order_is_accepted():
    b = FakeBlob(2, 2)
    t = lambda v: torch.tensor(v, dtype=torch.float64, requires_grad=True)
    G, E, A, B = t(np.ones(2)), t(np.ones(2)), t(np.ones(2)), t(np.ones(2))
    for _ in range(2):
        x = trajectory(b, G, E, A, B, np.zeros((2, 3)), 3)
        x.sum().backward()
    assert [c[0] for c in b.calls] == ["run", "adjoint", "run", "adjoint"]


def test_mixed_adjoint_without_refinement_warns():
    """GpuBlob.adjoint warns for an FP32-mixed run without FP64 refinement only"""
    from ffea_ci_demo_amd.blob import mixed_adjoint_warning
    assert "set_refinement" in mixed_adjoint_warning("mixed", 0.0)
    assert mixed_adjoint_warning("mixed", 1e-10) is None
    assert mixed_adjoint_warning("fp64", 0.0) is None
