"""MI355X-native VEGAS+ adaptive Monte Carlo integrator.

Drop-in for the hot path of emiliantolo/cuvegas: the integrand-definition
interface (HIP C++ source compiled by hipRTC), the parameters nbins, nstrat,
alpha, beta, neval, niter and the ``integrate`` entry point.
"""
from ._lib import CuvegasError, lib  # noqa: F401  (raises if the library is not built)
from .api import Integrator, Program, combine, default_n_strat, integrate  # noqa: F401
from . import integrands  # noqa: F401

__all__ = ["integrate", "Integrator", "Program", "combine", "default_n_strat", "integrands",
           "CuvegasError"]
