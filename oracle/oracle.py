"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product package (cuvegas_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

F_ONE, F_LINEAR, F_POLY, F_GAUSS, F_DGAUSS, F_OSC, F_CORNER, F_MIXTURE, F_GAUSSC = range(9)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


if not os.path.exists(LIB):
    build()
_lib = C.CDLL(LIB)

_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)
_PU = C.POINTER(C.c_uint32)
_PLL = C.POINTER(C.c_longlong)


class OrcParams(C.Structure):
    _fields_ = [
        ("d", C.c_int), ("nbins", C.c_int), ("nstrat", C.c_int), ("max_it", C.c_int),
        ("skip", C.c_int), ("K", C.c_int), ("n_eval_it", C.c_longlong), ("alpha", C.c_double),
        ("beta", C.c_double), ("seed", C.c_uint32), ("stream", C.c_uint32),
        ("precision", C.c_int), ("integrand", C.c_int), ("iparams", _PD), ("n_iparams", C.c_int),
        ("domain", _PD), ("domain_vol", C.c_double), ("iters", C.c_int), ("threads", C.c_int),
    ]


class OrcOut(C.Structure):
    _fields_ = [
        ("hist_I", _PD), ("hist_S2", _PD), ("hist_n", _PLL), ("x_edges", _PD), ("nh", _PI),
        ("nh_first", _PI), ("x_first", _PD), ("dh_first", _PD), ("JF_first", _PD),
        ("JF2_first", _PD), ("w_first", _PD), ("c_first", _PU), ("result", C.c_double),
        ("error", C.c_double), ("chi2_dof", C.c_double), ("tot_nevals", C.c_longlong),
    ]


assert _lib.orc_sizeof_params() == C.sizeof(OrcParams)
assert _lib.orc_sizeof_out() == C.sizeof(OrcOut)

_lib.orc_philox.argtypes = [_PU, C.c_uint32, C.c_uint32, _PU]
_lib.orc_log.argtypes = [C.c_double]; _lib.orc_log.restype = C.c_double
_lib.orc_exp.argtypes = [C.c_double]; _lib.orc_exp.restype = C.c_double
_lib.orc_pow.argtypes = [C.c_double, C.c_double]; _lib.orc_pow.restype = C.c_double
_lib.orc_sat_int.argtypes = [C.c_double]; _lib.orc_sat_int.restype = C.c_int
_lib.orc_bin_of.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _PU, _PU,
                            C.POINTER(C.c_uint64)]
_lib.orc_tree_sum.argtypes = [_PD, C.c_longlong]; _lib.orc_tree_sum.restype = C.c_double
_lib.orc_cube_sums.argtypes = [C.c_longlong, _PLL, C.c_longlong, _PD, _PD, _PD, _PD]
_lib.orc_results.argtypes = [C.c_longlong, C.c_double, C.c_double, _PI, _PD, _PD, _PD, _PD]
_lib.orc_alloc.argtypes = [C.c_longlong, _PD, C.c_double, C.c_longlong, _PI]
_lib.orc_prefix.argtypes = [C.c_longlong, _PI, _PLL]; _lib.orc_prefix.restype = C.c_longlong
_lib.orc_refine_dim.argtypes = [C.c_int, C.c_double, _PD, _PU, _PD, _PD]
_lib.orc_run.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcOut)]
_lib.orc_grad_replay.argtypes = [C.POINTER(OrcParams), _PD, _PI, C.c_int, _PD, _PD]
_lib.orc_sample_dump.argtypes = [C.POINTER(OrcParams), _PD, _PI, C.c_int, C.c_longlong, _PLL,
                                 _PI, _PD, _PD]


def _d(a):
    return a.ctypes.data_as(_PD)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    _lib.orc_philox(c.ctypes.data_as(_PU), int(key[0]), int(key[1]), out.ctypes.data_as(_PU))
    return out


def log(x): return _lib.orc_log(float(x))
def exp(x): return _lib.orc_exp(float(x))
def pow_(x, a): return _lib.orc_pow(float(x), float(a))
def sat_int(x): return _lib.orc_sat_int(float(x))


def bin_of(r, kn, nbins, nstrat):
    i, rem, low = C.c_uint32(), C.c_uint32(), C.c_uint64()
    _lib.orc_bin_of(r, kn, nbins, nstrat, C.byref(i), C.byref(rem), C.byref(low))
    return i.value, rem.value, low.value


def tree_sum(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return _lib.orc_tree_sum(_d(a), a.size)


def cube_sums(nh, v1, v2, K):
    nh = np.ascontiguousarray(nh, dtype=np.int32)
    C_ = nh.size
    nh_sum = np.zeros(C_ + 1, dtype=np.int64)
    _lib.orc_prefix(C_, nh.ctypes.data_as(_PI), nh_sum.ctypes.data_as(_PLL))
    v1 = np.ascontiguousarray(v1, dtype=np.float64)
    v2 = np.ascontiguousarray(v2, dtype=np.float64)
    JF = np.zeros(C_)
    JF2 = np.zeros(C_)
    rc = _lib.orc_cube_sums(C_, nh_sum.ctypes.data_as(_PLL), K, _d(v1), _d(v2), _d(JF), _d(JF2))
    assert rc == 0
    return JF, JF2


def results(V, beta, nh, JF, JF2):
    nh = np.ascontiguousarray(nh, dtype=np.int32)
    JF = np.ascontiguousarray(JF, dtype=np.float64)
    JF2 = np.ascontiguousarray(JF2, dtype=np.float64)
    dh = np.zeros(nh.size)
    sums = np.zeros(4)
    _lib.orc_results(nh.size, V, beta, nh.ctypes.data_as(_PI), _d(JF), _d(JF2), _d(dh), _d(sums))
    return dh, sums


def alloc(dh, D, N):
    dh = np.ascontiguousarray(dh, dtype=np.float64)
    nh = np.zeros(dh.size, dtype=np.int32)
    _lib.orc_alloc(dh.size, _d(dh), float(D), int(N), nh.ctypes.data_as(_PI))
    return nh


def refine_dim(alpha, weights, counts, x_edges):
    w = np.ascontiguousarray(weights, dtype=np.float64)
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    x = np.ascontiguousarray(x_edges, dtype=np.float64)
    out = np.zeros_like(x)
    _lib.orc_refine_dim(w.size, float(alpha), _d(w), c.ctypes.data_as(_PU), _d(x), _d(out))
    return out


def refine(alpha, weights, counts, x_edges):
    return np.stack([refine_dim(alpha, weights[i], counts[i], x_edges[i])
                     for i in range(x_edges.shape[0])])


def make_params(d, nbins, nstrat, n_eval_it, iters, integrand, iparams=(), *, skip=5, K=16,
                alpha=0.5, beta=0.75, seed=0, stream=0, precision=0, domain=None, threads=1):
    p = OrcParams()
    keep = {}
    p.d, p.nbins, p.nstrat, p.max_it, p.skip, p.K = d, nbins, nstrat, iters, skip, K
    p.n_eval_it, p.alpha, p.beta, p.seed, p.stream = int(n_eval_it), alpha, beta, seed, stream
    p.precision, p.integrand = precision, integrand
    ip = np.ascontiguousarray(iparams if len(iparams) else [0.0], dtype=np.float64)
    keep["ip"] = ip
    p.iparams, p.n_iparams = _d(ip), len(iparams)
    if domain is not None:
        dom = np.ascontiguousarray(domain, dtype=np.float64)
        keep["dom"] = dom
        p.domain = _d(dom)
        p.domain_vol = float(np.prod(dom[d:]))
    p.iters = iters
    p.threads = threads
    return p, keep


def run(p, keep, n_cubes, snapshots=True):
    """Run the oracle integrator; returns a dict of results and (optionally)
    the iteration-0 snapshots."""
    it = p.iters
    d, n = p.d, p.nbins
    o = OrcOut()
    arrs = {
        "hist_I": np.zeros(it), "hist_S2": np.zeros(it), "hist_n": np.zeros(it, dtype=np.int64),
        "x_edges": np.zeros((d, n + 1)), "nh": np.zeros(n_cubes, dtype=np.int32),
    }
    if snapshots:
        arrs.update({
            "nh_first": np.zeros(n_cubes, dtype=np.int32), "x_first": np.zeros((d, n + 1)),
            "dh_first": np.zeros(n_cubes), "JF_first": np.zeros(n_cubes),
            "JF2_first": np.zeros(n_cubes), "w_first": np.zeros((d, n)),
            "c_first": np.zeros((d, n), dtype=np.uint32),
        })
    for k, a in arrs.items():
        ptr_t = dict(OrcOut._fields_)[k]
        setattr(o, k, a.ctypes.data_as(ptr_t))
    rc = _lib.orc_run(C.byref(p), C.byref(o))
    if rc != 0:
        raise RuntimeError(f"orc_run failed: {rc}")
    arrs.update(result=o.result, error=o.error, chi2_dof=o.chi2_dof, tot_nevals=o.tot_nevals)
    return arrs


def sample_dump(p, keep, x_edges, nh, it, n):
    x_edges = np.ascontiguousarray(x_edges, dtype=np.float64)
    nh = np.ascontiguousarray(nh, dtype=np.int32)
    cube = np.zeros(n, dtype=np.int64)
    ids = np.zeros((n, p.d), dtype=np.int32)
    x = np.zeros((n, p.d))
    jf = np.zeros(n)
    rc = _lib.orc_sample_dump(C.byref(p), _d(x_edges), nh.ctypes.data_as(_PI), it, n,
                              cube.ctypes.data_as(_PLL), ids.ctypes.data_as(_PI), _d(x), _d(jf))
    assert rc == 0
    return cube, ids, x, jf


def grad_replay(p, keep, x_edges, nh, it, theta):
    """One frozen iteration of the mixture: (I_k, sigma_k^2, dI/dtheta, var)."""
    x_edges = np.ascontiguousarray(x_edges, dtype=np.float64)
    nh = np.ascontiguousarray(nh, dtype=np.int32)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    P = th.size
    out = np.zeros(2 + 2 * P)
    rc = _lib.orc_grad_replay(C.byref(p), _d(x_edges), nh.ctypes.data_as(_PI), it, _d(th), _d(out))
    assert rc == 0
    return out[0], out[1], out[2:2 + P], out[2 + P:]
