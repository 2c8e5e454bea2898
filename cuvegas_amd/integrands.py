"""Integrand library: HIP C++ sources for the sampling kernel, with closed forms.

Each builder returns an ``Integrand`` holding the source handed to hipRTC and,
where one exists, the exact integral over [0,1]^d.  Sources are precision
generic: they use ``vg_real`` (double in the FP64 form, float in the FP32 form)
and ``VG_NDIM`` (the dimension, fixed at compile time).  Parameters are baked
in as exact hexadecimal floating literals.

The set covers the reference's test integrands (src/commons/functions.cu:5-111:
linear, roosarnold, morokoffcalfisch, hellekalek, ridge, expon, cosin,
dirichlet, gaussian, sinexp, one) and the benchmark families (normalised
Gaussian peak, double Gaussian, oscillatory, corner peak, Gaussian mixture).
"""
from __future__ import annotations

import dataclasses
import itertools
import math
from typing import Callable, Optional, Sequence


@dataclasses.dataclass
class Integrand:
    name: str
    source: str
    n_dim: int
    exact: Optional[float] = None
    params: Sequence[float] = ()
    # parametrised form: theta vector and d exact / d theta (if known)
    n_theta: int = 0
    exact_grad: Optional[Callable] = None


def _lit(v: float) -> str:
    """Exact C++ literal of a double, cast to the working precision."""
    return f"((vg_real){float(v).hex()})"


def _arr(name: str, vals: Sequence[float]) -> str:
    body = ", ".join(float(v).hex() for v in vals)
    return f"__device__ static const double {name}[{len(vals)}] = {{{body}}};\n"


def _erf_mass(mu: float, s: float) -> float:
    """Integral of the normal density N(mu, s) over [0, 1]."""
    r = s * math.sqrt(2.0)
    return 0.5 * (math.erf((1.0 - mu) / r) + math.erf(mu / r))


# --- benchmark families --------------------------------------------------------

def gaussian_peak(n_dim: int, mu: float = 0.5, sigma: float = 0.1) -> Integrand:
    """prod_i N(x_i; mu, sigma): norm * exp(m * sum (x_i - mu)^2), m = -1 / (2 sigma^2)."""
    norm = (1.0 / (sigma * math.sqrt(2.0 * math.pi))) ** n_dim
    m = -1.0 / (2.0 * sigma * sigma)
    src = f"""
__device__ vg_real integrand(vg_real* x) {{
    vg_real q = (vg_real)0;
    for (int i = 0; i < VG_NDIM; ++i) {{ const vg_real a = x[i] - {_lit(mu)}; q = fma(a, a, q); }}
    return {_lit(norm)} * vg_expr(q * {_lit(m)});
}}
"""
    return Integrand("gaussian_peak", src, n_dim, _erf_mass(mu, sigma) ** n_dim, (mu, sigma, norm, m))


def double_gaussian(n_dim: int, c1: float = 0.33, c2: float = 0.67, sigma: float = 0.05) -> Integrand:
    """1/2 [g(x; c1) + g(x; c2)], g a normalised isotropic Gaussian of width sigma:
    g = norm * exp(m * |x - c|^2), m = -1 / (2 sigma^2)."""
    norm = (1.0 / (sigma * math.sqrt(2.0 * math.pi))) ** n_dim
    m = -1.0 / (2.0 * sigma * sigma)
    src = f"""
__device__ vg_real integrand(vg_real* x) {{
    vg_real q1 = (vg_real)0, q2 = (vg_real)0;
    for (int i = 0; i < VG_NDIM; ++i) {{
        const vg_real a = x[i] - {_lit(c1)}, b = x[i] - {_lit(c2)};
        q1 = fma(a, a, q1);
        q2 = fma(b, b, q2);
    }}
    return (vg_real)0.5 * {_lit(norm)} * (vg_expr(q1 * {_lit(m)}) + vg_expr(q2 * {_lit(m)}));
}}
"""
    exact = 0.5 * (_erf_mass(c1, sigma) ** n_dim + _erf_mass(c2, sigma) ** n_dim)
    return Integrand("double_gaussian", src, n_dim, exact, (c1, c2, sigma, norm, m))


def oscillatory(n_dim: int, u: float, a: Sequence[float]) -> Integrand:
    """cos(2 pi u + sum a_i x_i); exact = cos(2 pi u + sum a_i / 2) prod sin(a_i/2)/(a_i/2)."""
    a = [float(v) for v in a]
    assert len(a) == n_dim
    src = _arr("vg_osc_a", a) + f"""
__device__ vg_real integrand(vg_real* x) {{
    vg_real t = (vg_real)2 * (vg_real)3.14159265358979323846 * {_lit(u)};
    for (int i = 0; i < VG_NDIM; ++i) t = t + (vg_real)vg_osc_a[i] * x[i];
    return cos(t);
}}
"""
    exact = math.cos(2.0 * math.pi * u + 0.5 * sum(a))
    for ai in a:
        exact *= math.sin(ai / 2.0) / (ai / 2.0) if ai != 0 else 1.0
    return Integrand("oscillatory", src, n_dim, exact, (u, *a))


def corner_peak(n_dim: int, a: Sequence[float]) -> Integrand:
    """(1 + sum a_i x_i)^-(d+1); exact by inclusion-exclusion over the corners."""
    a = [float(v) for v in a]
    assert len(a) == n_dim
    src = _arr("vg_cp_a", a) + f"""
__device__ vg_real integrand(vg_real* x) {{
    vg_real t = (vg_real)1;
    for (int i = 0; i < VG_NDIM; ++i) t = t + (vg_real)vg_cp_a[i] * x[i];
    const vg_real r = (vg_real)1 / t;
    vg_real f = (vg_real)1;
    for (int i = 0; i <= VG_NDIM; ++i) f = f * r;
    return f;
}}
"""
    exact = None
    if n_dim <= 16:
        acc = 0.0
        for k in range(n_dim + 1):
            for S in itertools.combinations(range(n_dim), k):
                acc += (-1) ** k / (1.0 + sum(a[i] for i in S))
        exact = acc / (math.factorial(n_dim) * math.prod(a))
    return Integrand("corner_peak", src, n_dim, exact, tuple(a))


_MIXTURE_SRC = """
#define VG_MIX_K {K}
template <bool G>
__device__ inline vg_real vg_mixture(const vg_real* x, const vg_real* th, vg_real* gr) {{
    const vg_real c = sqrt((vg_real)2 * (vg_real)3.14159265358979323846);
    vg_real f = (vg_real)0;
    for (int k = 0; k < VG_MIX_K; ++k) {{
        const vg_real w = th[k], s = th[VG_MIX_K + k];
        const vg_real* mu = th + 2 * VG_MIX_K + k * VG_NDIM;
        // functions of theta only (the same for every sample): 1/s^2, 1/s, 1/norm
        const vg_real is2 = (vg_real)1 / (s * s), is = (vg_real)1 / s, m = (vg_real)-0.5 * is2;
        vg_real norm = (vg_real)1;
        for (int i = 0; i < VG_NDIM; ++i) norm = norm * (s * c);
        const vg_real inorm = (vg_real)1 / norm;
        vg_real q = (vg_real)0;
        for (int i = 0; i < VG_NDIM; ++i) {{ const vg_real a = x[i] - mu[i]; q = fma(a, a, q); }}
        const vg_real pk = vg_expr(q * m) * inorm;
        f = f + w * pk;
        if (G) {{
            const vg_real wp = w * pk;
            gr[k] = pk;
            gr[VG_MIX_K + k] = wp * (q * is2 - (vg_real)VG_NDIM) * is;
            for (int i = 0; i < VG_NDIM; ++i) gr[2 * VG_MIX_K + k * VG_NDIM + i] = wp * (x[i] - mu[i]) * is2;
        }}
    }}
    return f;
}}
"""


def mixture_exact(n_dim: int, theta: Sequence[float], K: int = 4):
    """Closed-form integral over [0,1]^d of sum_k w_k prod_i N(x_i; mu_ki, s_k) and its
    gradient with respect to theta = [w (K), s (K), mu (K x d)], from the normal CDF
    (erf) and the normal density at the faces."""
    th = [float(v) for v in theta]
    w, s = th[:K], th[K:2 * K]
    mu = [th[2 * K + k * n_dim: 2 * K + (k + 1) * n_dim] for k in range(K)]
    pdf = lambda x, m, sd: math.exp(-0.5 * ((x - m) / sd) ** 2) / (sd * math.sqrt(2 * math.pi))
    val = 0.0
    grad = [0.0] * (2 * K + K * n_dim)
    for k in range(K):
        m = [_erf_mass(mu[k][i], s[k]) for i in range(n_dim)]
        dm_dmu = [pdf(0.0, mu[k][i], s[k]) - pdf(1.0, mu[k][i], s[k]) for i in range(n_dim)]
        dm_ds = [-((1.0 - mu[k][i]) * pdf(1.0, mu[k][i], s[k]) + mu[k][i] * pdf(0.0, mu[k][i], s[k])) / s[k]
                 for i in range(n_dim)]
        prod = math.prod(m)
        val += w[k] * prod
        grad[k] = prod
        for i in range(n_dim):
            others = math.prod(m[j] for j in range(n_dim) if j != i)
            grad[2 * K + k * n_dim + i] = w[k] * others * dm_dmu[i]
            grad[K + k] += w[k] * others * dm_ds[i]
    return val, grad


def mixture_theta(n_dim: int, seed: int = 2, K: int = 4):
    """Parameters of the differentiable mixture workload: mu ~ U(0.2, 0.8),
    s ~ U(0.03, 0.15), w ~ U(0.5, 1.5) (numpy default_rng(seed))."""
    import numpy as np
    rng = np.random.default_rng(seed)
    mu = rng.uniform(0.2, 0.8, (K, n_dim))
    s = rng.uniform(0.03, 0.15, K)
    w = rng.uniform(0.5, 1.5, K)
    return [*w.tolist(), *s.tolist(), *mu.ravel().tolist()]


def gaussian_mixture_param(n_dim: int, theta: Sequence[float], K: int = 4) -> Integrand:
    """Parametrised mixture f(x; theta) = sum_k w_k prod_i N(x_i; mu_ki, s_k),
    theta = [w (K), s (K), mu (K x d)], with integrand_grad for the backward."""
    src = _MIXTURE_SRC.format(K=K) + """
__device__ vg_real integrand(vg_real* x, const vg_real* th) { return vg_mixture<false>(x, th, nullptr); }
__device__ vg_real integrand_grad(vg_real* x, const vg_real* th, vg_real* g) { return vg_mixture<true>(x, th, g); }
"""
    val, grad = mixture_exact(n_dim, theta, K)
    itg = Integrand("gaussian_mixture", src, n_dim, val, tuple(float(t) for t in theta),
                    n_theta=len(theta))
    itg.exact_grad = lambda th, d=n_dim, K=K: mixture_exact(d, th, K)[1]
    return itg


def gaussian_mixture(n_dim: int, w: Sequence[float], s: Sequence[float],
                     mu: Sequence[Sequence[float]]) -> Integrand:
    """sum_k w_k prod_i N(x_i; mu_ki, s_k) with the parameters compiled in."""
    K = len(w)
    flat = [float(v) for v in w] + [float(v) for v in s] + [float(m) for row in mu for m in row]
    src = _arr("vg_mix_p", flat) + _MIXTURE_SRC.format(K=K) + f"""
__device__ vg_real integrand(vg_real* x) {{
    vg_real th[{len(flat)}];
    for (int i = 0; i < {len(flat)}; ++i) th[i] = (vg_real)vg_mix_p[i];
    return vg_mixture<false>(x, th, nullptr);
}}
"""
    val, _ = mixture_exact(n_dim, flat, K)
    return Integrand("gaussian_mixture", src, n_dim, val, tuple(flat))


# --- polynomial integrands (bit-exact parity: no transcendental functions) ----

def one(n_dim: int) -> Integrand:
    src = "__device__ vg_real integrand(vg_real* x) { return (vg_real)1; }\n"
    return Integrand("one", src, n_dim, 1.0)


def linear(n_dim: int) -> Integrand:
    src = """
__device__ vg_real integrand(vg_real* x) {
    vg_real f = (vg_real)0;
    for (int i = 0; i < VG_NDIM; ++i) f = f + x[i];
    return f;
}
"""
    return Integrand("linear", src, n_dim, n_dim / 2.0)


def poly_peak(n_dim: int) -> Integrand:
    """prod_i 6 x_i (1 - x_i), integral 1."""
    src = """
__device__ vg_real integrand(vg_real* x) {
    vg_real f = (vg_real)1;
    for (int i = 0; i < VG_NDIM; ++i) f = f * ((vg_real)6 * x[i] * ((vg_real)1 - x[i]));
    return f;
}
"""
    return Integrand("poly_peak", src, n_dim, 1.0)


# --- the reference's integrand library (src/commons/functions.cu) -------------

def roosarnold(n_dim: int) -> Integrand:
    """(1/d) sum |4 x_i - 2|; each term through the single-precision fabsf, as
    the reference writes it (functions.cu:15-23)."""
    src = """
__device__ vg_real integrand(vg_real* x) {
    vg_real aux = (vg_real)0;
    for (int i = 0; i < VG_NDIM; ++i) aux = aux + (vg_real)fabsf((float)((vg_real)4 * x[i] - (vg_real)2));
    return aux / (vg_real)VG_NDIM;
}
"""
    return Integrand("roosarnold", src, n_dim, 1.0)


def morokoffcalfisch(n_dim: int) -> Integrand:
    src = """
__device__ vg_real integrand(vg_real* x) {
    vg_real f = (vg_real)1;
    for (int i = 0; i < VG_NDIM; ++i) f = f * pow(x[i], (vg_real)1 / (vg_real)VG_NDIM);
    return f * pow((vg_real)1 + (vg_real)1 / (vg_real)VG_NDIM, (vg_real)VG_NDIM);
}
"""
    return Integrand("morokoffcalfisch", src, n_dim, 1.0)


def hellekalek(n_dim: int) -> Integrand:
    """prod (x_i - 1/2) / sqrt(12), the root taken in single precision as the
    reference does (functions.cu:35-43)."""
    src = """
__device__ vg_real integrand(vg_real* x) {
    vg_real f = (vg_real)1;
    for (int i = 0; i < VG_NDIM; ++i) f = f * ((x[i] - (vg_real)0.5) / (vg_real)sqrtf(12.0f));
    return f;
}
"""
    return Integrand("hellekalek", src, n_dim, 0.0)


def ridge(n_dim: int) -> Integrand:
    """Sum of 1000 Gaussians along the diagonal (the reference's default integrand)."""
    src = """
__device__ vg_real integrand(vg_real* x) {
    const int N = 1000;
    const vg_real pi = (vg_real)3.14159265358979323846;
    const vg_real norm = ((vg_real)100 / pi) * ((vg_real)100 / pi) / (vg_real)N;
    vg_real f = (vg_real)0;
    for (int i = 0; i < N; ++i) {
        const vg_real x0 = (vg_real)i / ((vg_real)N - (vg_real)1);
        vg_real dx2 = (vg_real)0;
        for (int j = 0; j < VG_NDIM; ++j) { const vg_real a = x[j] - x0; dx2 = dx2 + a * a; }
        f = f + exp((vg_real)-100 * dx2);
    }
    return f * norm;
}
"""
    exact = None
    if n_dim == 4:
        N = 1000
        acc = 0.0
        for i in range(N):
            x0 = i / (N - 1.0)
            acc += (math.sqrt(math.pi / 100.0) * 0.5 * (math.erf(10 * (1 - x0)) + math.erf(10 * x0))) ** n_dim
        exact = (100.0 / math.pi) ** 2 / N * acc
    return Integrand("ridge", src, n_dim, exact)


def expon(n_dim: int) -> Integrand:
    src = """
__device__ vg_real integrand(vg_real* x) {
    vg_real f = (vg_real)0;
    for (int i = 0; i < VG_NDIM; ++i) f = f + x[i] * x[i];
    return exp(-f);
}
"""
    return Integrand("expon", src, n_dim, (math.sqrt(math.pi) / 2.0 * math.erf(1.0)) ** n_dim)


def cosin(n_dim: int) -> Integrand:
    src = """
__device__ vg_real integrand(vg_real* x) {
    vg_real f = (vg_real)1;
    for (int i = 0; i < VG_NDIM; ++i) f = f * cos(x[i]);
    return f;
}
"""
    return Integrand("cosin", src, n_dim, math.sin(1.0) ** n_dim)


def dirichlet(n_dim: int) -> Integrand:
    """prod x_i^4 normalised by Gamma(5)^d / Gamma(5d) (reference normalisation)."""
    b = math.gamma(5.0) ** n_dim / math.gamma(5.0 * n_dim)
    src = f"""
__device__ vg_real integrand(vg_real* x) {{
    vg_real f = (vg_real)1;
    for (int i = 0; i < VG_NDIM; ++i) f = f * (x[i] * x[i] * x[i] * x[i]);
    return f / {_lit(b)};
}}
"""
    return Integrand("dirichlet", src, n_dim, (1.0 / 5.0) ** n_dim / b)


def gaussian(n_dim: int) -> Integrand:
    """The reference test Gaussian: sigma^2 = 1e-4, mu = 0.5 (vegas_cuda/test.py:16-23)."""
    sigma2 = 1e-4
    norm = 1.0 / (2.0 * math.pi * sigma2) ** (n_dim / 2.0)
    src = f"""
__device__ vg_real integrand(vg_real* x) {{
    vg_real f = (vg_real)0;
    for (int i = 0; i < VG_NDIM; ++i) {{ const vg_real a = x[i] - (vg_real)0.5; f = f + a * a; }}
    return exp((vg_real)-0.5 * f / {_lit(sigma2)}) * {_lit(norm)};
}}
"""
    return Integrand("gaussian", src, n_dim, _erf_mass(0.5, math.sqrt(sigma2)) ** n_dim)


def sinexp(n_dim: int) -> Integrand:
    src = """
__device__ vg_real integrand(vg_real* x) { return sin(x[0]) + exp(x[1]); }
"""
    return Integrand("sinexp", src, n_dim, (1.0 - math.cos(1.0)) + (math.e - 1.0))


LIBRARY = {
    "one": one, "linear": linear, "poly_peak": poly_peak, "roosarnold": roosarnold,
    "morokoffcalfisch": morokoffcalfisch, "hellekalek": hellekalek, "ridge": ridge,
    "expon": expon, "cosin": cosin, "dirichlet": dirichlet, "gaussian": gaussian,
    "sinexp": sinexp, "gaussian_peak": gaussian_peak, "double_gaussian": double_gaussian,
}


def by_name(name: str, n_dim: int) -> Integrand:
    if name not in LIBRARY:
        raise KeyError(f"unknown integrand {name!r}; known: {sorted(LIBRARY)}")
    return LIBRARY[name](n_dim)
