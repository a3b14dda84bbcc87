"""Host logic of the torch.autograd bridge (fermiphasespace_amd.autograd) with a stand-in adjoint.

The stand-in returns fixed gradients, so the test checks the plumbing: J is the adjoint's objective,
backward scales the stored gradients by grad_output, and a run without a pulse gives no knot gradient.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from fermiphasespace_amd import autograd as ag  # noqa: E402


class StandIn:
    def __init__(self, L, n):
        self.calls = []
        self.L, self.n = L, n

    def adjoint(self, U, tM, knots):
        self.calls.append((U.copy(), tM.copy(), None if knots is None else knots.copy()))
        return {"J": float(U.sum() + 0.5 * tM.sum()),
                "grad_u": np.arange(1.0, self.L + 1),
                "grad_t": np.full((self.L, self.L), 0.25),
                "grad_pulse": None if knots is None else np.linspace(-1.0, 1.0, self.n)}


def test_forward_backward_with_knots():
    L, n = 4, 3
    U = torch.full((L,), 2.0, dtype=torch.float64, requires_grad=True)
    tM = torch.zeros((L, L), dtype=torch.float64, requires_grad=True)
    knots = torch.zeros(n, dtype=torch.float64, requires_grad=True)
    prob = StandIn(L, n)
    J = ag.double_occupancy(U, tM, knots, prob)
    assert J.item() == pytest.approx(8.0)
    (3.0 * J).backward()
    assert torch.allclose(U.grad, 3.0 * torch.arange(1.0, L + 1, dtype=torch.float64))
    assert torch.allclose(tM.grad, torch.full((L, L), 0.75, dtype=torch.float64))
    assert torch.allclose(knots.grad, 3.0 * torch.linspace(-1.0, 1.0, n, dtype=torch.float64))
    assert len(prob.calls) == 1  # backward reuses the forward's adjoint run


def test_without_pulse():
    L = 3
    U = torch.ones(L, dtype=torch.float64, requires_grad=True)
    tM = torch.ones((L, L), dtype=torch.float64, requires_grad=True)
    prob = StandIn(L, 0)
    J = ag.double_occupancy(U, tM, None, prob)
    J.backward()
    assert prob.calls[0][2] is None
    assert torch.allclose(U.grad, torch.arange(1.0, L + 1, dtype=torch.float64))
