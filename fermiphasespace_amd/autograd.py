"""torch.autograd bridge to the discrete adjoint of the fTWA engine (fpsftwa_run_adjoint).

    J = double_occupancy(U, tM, knots, problem)
    J.backward()            # U.grad, tM.grad, knots.grad from one adjoint run

J is the ensemble double occupancy at the last time point, mean over sites of NuNd
(src/FHfTWASolver.cpp:395-398, :524-531).  The forward already runs the adjoint (the gradient costs one
checkpointed reverse sweep on top of the forward integration), so backward only scales the stored
gradients by the incoming grad_output.  tM.grad is the derivative with respect to each entry of the
hopping pattern (tM_kl != 0) and the diagonal, each entry independent; other entries get 0.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass
class FtwaProblem:
    """Everything of a run except the differentiated parameters."""
    n_up: np.ndarray
    n_dn: np.ndarray
    dt: float = 0.01
    t_max: float = 1.0
    n_sub: int = 1
    n_traj: int = 1024
    seed: int = 0
    engine: str = "fp32"
    checkpoint_every: int = 50
    pot: np.ndarray | None = None
    pol: tuple = (0.0, 0.0, 0.0)
    disp: np.ndarray | None = None
    rot_up: np.ndarray | None = None
    rot_dn: np.ndarray | None = None
    _last: dict = field(default_factory=dict, repr=False)

    def adjoint(self, U: np.ndarray, tM: np.ndarray, knots: np.ndarray | None) -> dict:
        """One adjoint run at (U, tM, knots): {"J", "grad_u", "grad_t", "grad_pulse"}."""
        pulse = {}
        if knots is not None and np.size(knots):
            pulse = dict(pulse_type=_lib.PULSE_KNOTS, knots=np.asarray(knots, dtype=np.float64), pol=self.pol,
                         disp=self.disp)
        with _lib.Engine(np.asarray(tM, dtype=np.float64), coulomb_type=_lib.COULOMB_ONSITE,
                         U=np.diag(np.asarray(U, dtype=np.float64)), pot=self.pot, **pulse) as eng:
            eng.set_initial_state(self.n_up, self.n_dn, self.rot_up, self.rot_dn)
            eng.set_params(self.dt, self.t_max, self.n_sub, self.n_traj, seed=self.seed, engine=self.engine)
            res = eng.adjoint(self.checkpoint_every)
        self._last = res
        return res


def _torch():
    import torch  # plumbing only: the arithmetic runs in the engine

    return torch


def _make_function():
    torch = _torch()

    class DoubleOccupancy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, U, tM, knots, problem):
            k = knots.detach().cpu().numpy() if knots is not None and knots.numel() else None
            res = problem.adjoint(U.detach().cpu().numpy(), tM.detach().cpu().numpy(), k)
            gp = res["grad_pulse"] if res["grad_pulse"] is not None else np.zeros(0)
            ctx.save_for_backward(torch.as_tensor(res["grad_u"], dtype=U.dtype),
                                  torch.as_tensor(res["grad_t"], dtype=tM.dtype),
                                  torch.as_tensor(gp, dtype=U.dtype))
            ctx.has_knots = knots is not None
            return torch.tensor(res["J"], dtype=U.dtype)

        @staticmethod
        def backward(ctx, g):
            gu, gt, gp = ctx.saved_tensors
            return g * gu, g * gt, (g * gp if ctx.has_knots else None), None

    return DoubleOccupancy


_FN = None


def double_occupancy(U, tM, knots, problem: FtwaProblem):
    """Differentiable J(U [L], tM [L, L], knots [n] or None) for the fixed inputs in problem."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(U, tM, knots, problem)
