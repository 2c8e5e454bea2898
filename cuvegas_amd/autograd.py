"""torch.autograd binding of the VEGAS+ integral of a parametrised integrand.

Forward: the adaptive integrator runs with the integrand's parameters theta and
records the map and allocation of every averaged iteration.  Backward: a
hand-written replay kernel regenerates each recorded iteration's samples from
the same Philox counters (nothing per sample is stored) and returns
d I / d theta of the weighted mean, with the forward's inverse-variance weights
held fixed.  The integrand source must define

    vg_real integrand(vg_real* x, const vg_real* theta);
    vg_real integrand_grad(vg_real* x, const vg_real* theta, vg_real* grad);

(``grad`` receives df/dtheta; see cuvegas_amd.integrands.gaussian_mixture_param).
PyTorch is used only for the autograd plumbing; theta is copied to the host.
"""
from __future__ import annotations

import weakref
from typing import Optional

import numpy as np
import torch

from .api import Integrator


class _VegasIntegralFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, module: "DifferentiableIntegral"):
        th = theta.detach().to("cpu", torch.float64).numpy()
        integ = module._new_state(th)
        integ.run(module.max_it, use_graph=module.use_graph)
        r = integ.result()
        module.last = {"result": r["result"], "error": r["error"], "chi2_dof": r["chi2_dof"],
                       "tot_nevals": r["tot_nevals"]}
        ctx.integ = integ
        ctx.module = module
        module._pending.add(integ)
        return theta.new_tensor(r["result"])

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        integ, module = ctx.integ, ctx.module
        _, _, g, gv = integ.backward()
        module._pending.discard(integ)
        module.last_grad_var = gv
        grad = torch.as_tensor(g, dtype=grad_out.dtype, device=grad_out.device)
        return grad_out * grad, None


class DifferentiableIntegral:
    """I(theta) = integral over [0,1]^d (or a box) of f(x; theta), differentiable in theta.

    Calling the object with a tensor theta runs a fresh adaptive integration
    and returns a 0-d tensor carrying a grad_fn.
    """

    def __init__(self, source: str, n_dim: int, n_theta: int, n_eval_it: int, *,
                 max_it: int = 10, skip: int = 5, n_intervals: int = 1024,
                 n_strat: Optional[int] = None, alpha: float = 0.5, beta: float = 0.75,
                 seed: int = 0, precision: str = "fp64", domain=None, use_graph: bool = True,
                 options: Optional[str] = None):
        self.source, self.n_dim, self.n_theta, self.n_eval_it = source, n_dim, n_theta, n_eval_it
        self.kw = dict(n_intervals=n_intervals, n_strat=n_strat, max_it=max_it, skip=skip,
                       alpha=alpha, beta=beta, seed=seed, precision=precision, domain=domain)
        self.max_it = max_it
        self.use_graph = use_graph
        self.options = ("grad " + (options or "")).strip()
        self._program = None
        self.last = None
        self.last_grad_var = None
        self.state: Optional[Integrator] = None
        # integrators of forwards whose backward has not run yet
        self._pending: "weakref.WeakSet[Integrator]" = weakref.WeakSet()

    def _new_state(self, theta: np.ndarray) -> Integrator:
        # the previous forward's device buffers are released now, unless a
        # graph whose backward has not run yet still needs them (it replays on
        # that integrator, which is freed with the graph)
        if self.state is not None:
            old, self.state = self.state, None
            if old not in self._pending:
                old.close()
        integ = Integrator(self.source, self.n_dim, self.n_eval_it, theta=theta,
                           options=self.options, program=self._program, **self.kw)
        self._program = integ.program
        integ.set_record(True)
        self.state = integ
        return integ

    def __call__(self, theta: torch.Tensor) -> torch.Tensor:
        if theta.numel() != self.n_theta:
            raise ValueError(f"theta has {theta.numel()} entries, expected {self.n_theta}")
        return _VegasIntegralFn.apply(theta, self)

    def replay(self, theta) -> float:
        """Value of the frozen estimator (last forward's maps, allocations,
        counters and weights) at another theta: the function whose derivative
        backward() returns; use it for finite differences."""
        if self.state is None:
            raise RuntimeError("run a forward pass first")
        v, _, _, _ = self.state.backward(np.asarray(theta, dtype=np.float64))
        return v
