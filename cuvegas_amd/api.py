"""Python host API over the C ABI: programs, integrator states, integrate().

``integrate`` mirrors the reference entry point
(vegas_cuda/src/vegas_cuda_extension/python_wrapper.py:9-77): same arguments,
same defaults, same ``(result, dict)`` return with the reference's 13 keys,
plus extra keys (chi2_dof, per-iteration estimates, evals_per_s, precision).
``device_function`` is a HIP C++ source string defining ``integrand`` (or the
name of a library integrand, see cuvegas_amd.integrands), compiled by hipRTC;
the reference takes a numba device function compiled to PTX.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import threading
import time
from collections import OrderedDict
from typing import Optional, Sequence

import numpy as np

from . import integrands as _ints
from ._lib import (CVG_FP32C, CVG_FP64, CuvegasError, check, cvg_params, cvg_result, lib)

STATE_FORMAT = "cuvegas_amd.state.v1"

PRECISIONS = {"fp64": CVG_FP64, "fp32": CVG_FP32C, "fp32c": CVG_FP32C}


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _uptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint))


def _llptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_longlong))


def default_n_strat(n_evals: int, max_it: int, n_dim: int) -> int:
    """Reference default (python_wrapper.py:15-18): int((n_eval_it / 4)^(1/d))."""
    n_eval_it = n_evals // max_it
    return int(math.pow((n_eval_it / 4.0), 1.0 / n_dim))


def make_params(n_dim: int, n_intervals: int, n_strat: int, max_it: int, skip: int,
                n_eval_it: int, alpha: float, beta: float, seed: int, precision="fp64",
                stream_id: int = 0, run_len: int = 0, domain: Optional[Sequence] = None,
                n_theta: int = 0):
    p = cvg_params()
    p.n_dim = n_dim
    p.n_intervals = n_intervals
    p.n_strat = n_strat
    p.max_it = max_it
    p.skip = skip
    p.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    p.n_eval_it = int(n_eval_it)
    p.alpha = alpha
    p.beta = beta
    p.seed = int(seed) & 0xFFFFFFFF
    p.stream_id = int(stream_id) & 0xFFFFFFFF
    p.run_len = run_len
    p.n_theta = n_theta
    keep = None
    if domain is not None:
        keep = np.ascontiguousarray(domain, dtype=np.float64)
        p.has_domain = 1
        p.domain = _dptr(keep)
    return p, keep


class Program:
    """hipRTC-compiled sampling kernel for one integrand source and shape."""

    def __init__(self, source: str, params: cvg_params, options: Optional[str] = None):
        h = C.c_void_p()
        rc = lib.cvg_program_create(source.encode(), C.byref(params),
                                    options.encode() if options else None, C.byref(h))
        self._h = h
        if rc < 0:
            log = lib.cvg_program_log(h).decode(errors="replace") if h.value else ""
            msg = lib.cvg_last_error().decode(errors="replace")
            if h.value:
                lib.cvg_program_destroy(h)
                self._h = C.c_void_p()
            raise CuvegasError(rc, msg + ("\n" + log if log and log not in msg else ""))
        self.log = lib.cvg_program_log(h).decode(errors="replace")
        self.code_size = lib.cvg_program_code_size(h)

    def code(self) -> bytes:
        """The gfx950 code object (ELF) of the sampling kernel."""
        buf = C.create_string_buffer(int(self.code_size))
        n = lib.cvg_program_get_code(self._h, buf, int(self.code_size))
        return buf.raw[:n]

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h and self._h.value:
            lib.cvg_program_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Compiled programs kept for reuse within the process (least recently used
# dropped beyond the bound); a program can back any number of states.  The
# lock covers the table only: a compile runs outside it (ctypes releases the
# GIL), and two threads that miss on the same key both compile, the second
# result replacing the first in the table.
_PROGRAMS: "OrderedDict[tuple, Program]" = OrderedDict()
_PROGRAMS_MAX = 16
_PROGRAMS_LOCK = threading.Lock()


def cached_program(source: str, params: cvg_params, options: Optional[str] = None) -> Program:
    """The hipRTC program of (source, compile-time shape, options), compiled on
    first use and reused afterwards (saves the compile, ~0.2 s, on repeated calls)."""
    key = (source, params.n_dim, params.n_intervals, params.n_strat, params.precision, params.run_len,
           params.has_domain, params.n_theta, options or "")
    with _PROGRAMS_LOCK:
        prog = _PROGRAMS.get(key)
        if prog is not None:
            _PROGRAMS.move_to_end(key)
            return prog
    prog = Program(source, params, options)
    with _PROGRAMS_LOCK:
        _PROGRAMS[key] = prog
        while len(_PROGRAMS) > _PROGRAMS_MAX:
            _PROGRAMS.popitem(last=False)
    return prog


class Integrator:
    """One integral's device state and iteration loop."""

    def __init__(self, source: str, n_dim: int, n_eval_it: int, *, n_intervals: int = 1024,
                 n_strat: Optional[int] = None, max_it: int = 20, skip: int = 5,
                 alpha: float = 0.5, beta: float = 0.75, seed: int = 0, precision="fp64",
                 stream_id: int = 0, run_len: int = 0, domain: Optional[Sequence] = None,
                 options: Optional[str] = None, theta: Optional[Sequence[float]] = None,
                 program: Optional[Program] = None, cache_program: bool = False):
        if n_strat is None:
            n_strat = max(1, int(math.pow(n_eval_it / 4.0, 1.0 / n_dim)))
        n_theta = 0 if theta is None else len(theta)
        self.params, self._dom = make_params(n_dim, n_intervals, n_strat, max_it, skip, n_eval_it,
                                             alpha, beta, seed, precision, stream_id, run_len,
                                             domain, n_theta)
        if program is None:
            program = (cached_program(source, self.params, options) if cache_program
                       else Program(source, self.params, options))
        self.program = program
        h = C.c_void_p()
        rc = lib.cvg_state_create(self.program.handle, C.byref(self.params), C.byref(h))
        if rc < 0 and cache_program:
            # a cached program is bound to the device it was first loaded on:
            # compile a fresh one for this state
            self.program = Program(source, self.params, options)
            rc = lib.cvg_state_create(self.program.handle, C.byref(self.params), C.byref(h))
        check(rc)
        self._h = h
        self.n_dim, self.n_intervals, self.n_strat = n_dim, n_intervals, n_strat
        self.n_cubes = int(lib.cvg_state_n_cubes(h))
        self.max_it, self.skip = max_it, skip
        self.theta = None
        if theta is not None:
            self.set_theta(theta)

    # -- loop --------------------------------------------------------------
    def run(self, n_iter: int, use_graph: bool = True) -> "Integrator":
        check(lib.cvg_state_run(self._h, int(n_iter), 1 if use_graph else 0))
        return self

    def result(self) -> dict:
        r = cvg_result()
        check(lib.cvg_state_result(self._h, C.byref(r)))
        return {f: getattr(r, f) for f, _ in cvg_result._fields_}

    def history(self):
        n = self.max_it
        I = np.zeros(n)
        S2 = np.zeros(n)
        ne = np.zeros(n, dtype=np.int64)
        k = check(lib.cvg_state_history(self._h, _dptr(I), _dptr(S2), _llptr(ne), n))
        return I[:k], S2[:k], ne[:k]

    def history_unbiased(self) -> np.ndarray:
        """Per-iteration variance of I_k from the unbiased per-cube variances
        sigma_h^2 / (n_h - 1) (the reference's sigma_k^2 divides by n_h)."""
        S2u = np.zeros(self.max_it)
        k = check(lib.cvg_state_history_unbiased(self._h, _dptr(S2u), self.max_it))
        return S2u[:k]

    def set_timing(self, enable: bool = True) -> None:
        check(lib.cvg_state_set_timing(self._h, 1 if enable else 0))

    def timing(self):
        """Per-iteration device times (ms) of eager runs: (fill kernel, whole iteration)."""
        cap = 4096
        f = np.zeros(cap)
        t = np.zeros(cap)
        k = check(lib.cvg_state_get_timing(self._h, _dptr(f), _dptr(t), cap))
        k = min(k, cap)
        return f[:k], t[:k]

    # -- state I/O -----------------------------------------------------------
    def get_map(self) -> np.ndarray:
        x = np.zeros((self.n_dim, self.n_intervals + 1))
        check(lib.cvg_state_get_map(self._h, _dptr(x)))
        return x

    def set_map(self, x_edges) -> None:
        x = np.ascontiguousarray(x_edges, dtype=np.float64).reshape(self.n_dim, self.n_intervals + 1)
        check(lib.cvg_state_set_map(self._h, _dptr(x)))

    def get_nh(self) -> np.ndarray:
        nh = np.zeros(self.n_cubes, dtype=np.int32)
        check(lib.cvg_state_get_nh(self._h, _iptr(nh)))
        return nh

    def set_nh(self, nh) -> None:
        a = np.ascontiguousarray(nh, dtype=np.int32)
        check(lib.cvg_state_set_nh(self._h, _iptr(a)))

    def get_dh(self) -> np.ndarray:
        dh = np.zeros(self.n_cubes)
        check(lib.cvg_state_get_dh(self._h, _dptr(dh)))
        return dh

    def get_cube_sums(self):
        a = np.zeros(self.n_cubes)
        b = np.zeros(self.n_cubes)
        check(lib.cvg_state_get_cube_sums(self._h, _dptr(a), _dptr(b)))
        return a, b

    def get_bins(self):
        w = np.zeros((self.n_dim, self.n_intervals))
        c = np.zeros((self.n_dim, self.n_intervals), dtype=np.uint32)
        check(lib.cvg_state_get_bins(self._h, _dptr(w), _uptr(c)))
        return w, c

    def set_bins(self, weights, counts) -> None:
        w = np.ascontiguousarray(weights, dtype=np.float64)
        c = np.ascontiguousarray(counts, dtype=np.uint32)
        check(lib.cvg_state_set_bins(self._h, _dptr(w), _uptr(c)))

    def set_iteration(self, it: int) -> None:
        check(lib.cvg_state_set_iteration(self._h, int(it)))

    def set_history(self, I, sigma2, nevals, sigma2_unbiased=None) -> None:
        """Restore the per-iteration record; the next iteration is len(I)."""
        I = np.ascontiguousarray(I, dtype=np.float64)
        S2 = np.ascontiguousarray(sigma2, dtype=np.float64)
        ne = np.ascontiguousarray(nevals, dtype=np.int64)
        if not (I.shape == S2.shape == ne.shape and I.ndim == 1):
            raise ValueError("I, sigma2 and nevals must be 1-D arrays of one length")
        S2u = None
        if sigma2_unbiased is not None:
            S2u = np.ascontiguousarray(sigma2_unbiased, dtype=np.float64)
            if S2u.shape != I.shape:
                raise ValueError("sigma2_unbiased must have the length of I")
        check(lib.cvg_state_set_history(self._h, _dptr(I), _dptr(S2), _llptr(ne),
                                        None if S2u is None else _dptr(S2u), int(I.size)))

    # -- on-disk state -------------------------------------------------------
    # An .npz archive of plain arrays (np.load(path, allow_pickle=False) reads
    # it): the shape and run parameters, the map edges, the allocation n_h and
    # the per-iteration (I_k, sigma_k^2, samples), and the parameters theta of
    # a parametrised integrand.  Restoring it reproduces the iteration the
    # saved run would have run next: same map, same allocation and sample
    # layout, same Philox counter block, same parameters.  Not saved: the
    # compiled integrand (the caller passes the source again), the bin
    # accumulators (cleared at the end of every iteration) and the autograd
    # record.

    def save(self, path) -> None:
        """Write the run's state (see above) to `path` (.npz)."""
        I, S2, ne = self.history()
        p = self.params
        dom = self._dom if self._dom is not None else np.zeros(0)
        np.savez(path, format=np.array(STATE_FORMAT), n_dim=p.n_dim, n_intervals=p.n_intervals,
                 n_strat=p.n_strat, n_eval_it=p.n_eval_it, max_it=p.max_it, skip=p.skip,
                 alpha=p.alpha, beta=p.beta, seed=p.seed, stream_id=p.stream_id,
                 precision=p.precision, run_len=p.run_len, domain=dom, x_edges=self.get_map(),
                 nh=self.get_nh(), I=I, sigma2=S2, nevals=ne, sigma2_unbiased=self.history_unbiased(),
                 n_theta=p.n_theta,
                 theta=self.theta if self.theta is not None else np.zeros(0))

    def restore(self, path) -> "Integrator":
        """Load a saved state into this integrator (same n_dim, n_intervals, n_strat)."""
        with np.load(path, allow_pickle=False) as z:
            if "format" not in z or str(z["format"]) != STATE_FORMAT:
                raise ValueError(f"{path}: not a {STATE_FORMAT} archive")
            shape = (int(z["n_dim"]), int(z["n_intervals"]), int(z["n_strat"]))
            if shape != (self.n_dim, self.n_intervals, self.n_strat):
                raise ValueError(f"saved shape (n_dim, n_intervals, n_strat) = {shape} differs from "
                                 f"{(self.n_dim, self.n_intervals, self.n_strat)}")
            if z["I"].size > self.max_it:
                raise ValueError(f"saved run has {z['I'].size} iterations; max_it is {self.max_it}")
            n_theta = int(z["n_theta"]) if "n_theta" in z else 0
            if n_theta != self.params.n_theta:
                raise ValueError(f"saved integrand has {n_theta} parameters; this integrator has "
                                 f"{self.params.n_theta}")
            self.set_map(z["x_edges"])
            self.set_nh(z["nh"])
            self.set_history(z["I"], z["sigma2"], z["nevals"],
                             z["sigma2_unbiased"] if "sigma2_unbiased" in z else None)
            if n_theta > 0:
                self.set_theta(z["theta"])
        return self

    @classmethod
    def from_state(cls, path, source: str, *, max_it: Optional[int] = None,
                   options: Optional[str] = None, program: Optional[Program] = None) -> "Integrator":
        """A new integrator with the saved parameters, restored from `path`;
        max_it (history capacity) may be raised to continue the run."""
        with np.load(path, allow_pickle=False) as z:
            if "format" not in z or str(z["format"]) != STATE_FORMAT:
                raise ValueError(f"{path}: not a {STATE_FORMAT} archive")
            dom = z["domain"]
            kw = dict(n_intervals=int(z["n_intervals"]), n_strat=int(z["n_strat"]),
                      max_it=int(max_it if max_it is not None else z["max_it"]), skip=int(z["skip"]),
                      alpha=float(z["alpha"]), beta=float(z["beta"]), seed=int(z["seed"]),
                      precision=int(z["precision"]), stream_id=int(z["stream_id"]),
                      run_len=int(z["run_len"]), domain=dom.reshape(2, -1) if dom.size else None)
            n_dim, n_eval_it = int(z["n_dim"]), int(z["n_eval_it"])
            theta = z["theta"].copy() if "n_theta" in z and int(z["n_theta"]) > 0 else None
        itg = cls(source, n_dim, n_eval_it, options=options, program=program, theta=theta, **kw)
        return itg.restore(path)

    def set_theta(self, theta) -> None:
        t = np.ascontiguousarray(theta, dtype=np.float64).ravel()
        check(lib.cvg_state_set_theta(self._h, _dptr(t), int(t.size)))
        self.theta = t.copy()

    def set_stream(self, hip_stream: int) -> None:
        check(lib.cvg_state_set_stream(self._h, C.c_void_p(hip_stream)))

    # -- single steps ----------------------------------------------------------
    def fill(self) -> None:
        check(lib.cvg_state_fill(self._h))

    def refine(self) -> None:
        check(lib.cvg_state_refine(self._h))

    def reallocate(self, dh, D: float) -> None:
        a = np.ascontiguousarray(dh, dtype=np.float64)
        check(lib.cvg_state_reallocate(self._h, _dptr(a), float(D)))

    def dump(self, it: int, n: int):
        cube = np.zeros(n, dtype=np.int64)
        ids = np.zeros((n, self.n_dim), dtype=np.int32)
        x = np.zeros((n, self.n_dim))
        jf = np.zeros(n)
        check(lib.cvg_state_dump(self._h, int(it), int(n), _llptr(cube), _iptr(ids), _dptr(x), _dptr(jf)))
        return cube, ids, x, jf

    # -- differentiation -------------------------------------------------------
    def set_record(self, enable: bool = True) -> None:
        """Keep the map and allocation of every iteration >= skip for backward()."""
        check(lib.cvg_state_set_record(self._h, 1 if enable else 0))

    def snapshot(self, it: int):
        x = np.zeros((self.n_dim, self.n_intervals + 1))
        nh = np.zeros(self.n_cubes, dtype=np.int32)
        check(lib.cvg_state_get_snapshot(self._h, int(it), _dptr(x), _iptr(nh)))
        return x, nh

    def backward(self, theta=None):
        """Replay the recorded iterations: returns (value, value_var, grad, grad_var)
        of the frozen estimator at theta (default: the forward's parameters)."""
        P = self.params.n_theta
        g = np.zeros(P)
        gv = np.zeros(P)
        v = C.c_double()
        vv = C.c_double()
        th = None if theta is None else np.ascontiguousarray(theta, dtype=np.float64)
        check(lib.cvg_state_backward(self._h, None if th is None else _dptr(th), _dptr(g), _dptr(gv),
                                     C.byref(v), C.byref(vv)))
        return v.value, vv.value, g, gv

    def backward_timing(self) -> float:
        """Device time (ms) of the last backward()'s replay kernels."""
        ms = C.c_double()
        check(lib.cvg_state_backward_timing(self._h, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.cvg_state_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def combine(I: np.ndarray, S2: np.ndarray, skip: int):
    """Inverse-variance weighted mean of iterations >= skip, its error and chi2/dof
    (reference cuda_wrapper.cu:758-766 and :869; chi2/dof is an addition), by
    the rule cvg_state_result applies: iterations of zero variance are exact,
    and if there are any the result is their mean, with error and chi2/dof 0
    (the reference's sums give inf / inf = NaN there)."""
    I = np.asarray(I, dtype=np.float64)
    S2 = np.asarray(S2, dtype=np.float64)
    idx = list(range(skip, len(I)))
    exact = [i for i in idx if S2[i] == 0.0]
    if exact:
        return sum(I[i] for i in exact) / len(exact), 0.0, 0.0
    res, sig, K = 0.0, 0.0, 0
    for i in idx:
        res += I[i] / S2[i]
        sig += 1.0 / S2[i]
        K += 1
    if K > 0:
        res /= sig
    err = 1.0 / math.sqrt(sig) if sig > 0 else math.inf
    chi2 = sum((I[i] - res) ** 2 / S2[i] for i in idx)
    return res, err, (chi2 / (K - 1) if K > 1 else 0.0)


def integrate(device_function, n_dim, domain=None, n_evals=1000000, max_it=20, skip=5,
              batch_size=1048576, n_intervals=1024, n_strat=None, alpha=0.5, beta=0.75,
              seed=None, multi_gpu=None, multi_cpu=None, gpu_num=-1, cpu_num=-1, gpu_thr=0.1,
              *, precision="fp64", use_graph=True, run_len=0, options=None):
    """VEGAS+ integral of ``device_function`` over ``domain`` (default [0,1]^n_dim).

    Mirrors python_wrapper.integrate (python_wrapper.py:9-77).  ``n_evals`` is
    the total over all iterations (n_evals // max_it per iteration).
    Returns (result, info) where info holds the reference's keys error, n_strat,
    n_cubes, n_intervals, tot_nevals, tot_time, avg_it_time, seed, multi_gpu,
    multi_cpu, gpu_num, omp_threads_num, gpu_thr, plus chi2_dof, iterations,
    evals_per_s, precision and exact (library integrands).
    """
    t0 = time.perf_counter()
    if seed is None:  # python_wrapper.py:12-13
        seed = int.from_bytes(os.urandom(4), "big")
    if n_strat is None:  # python_wrapper.py:15-18
        n_strat = default_n_strat(n_evals, max_it, n_dim)
    exact = None
    if isinstance(device_function, _ints.Integrand):
        exact = device_function.exact
        source = device_function.source
    elif isinstance(device_function, str) and device_function in _ints.LIBRARY:
        lib_int = _ints.by_name(device_function, n_dim)
        exact, source = lib_int.exact, lib_int.source
    elif isinstance(device_function, str):
        source = device_function
    else:
        raise TypeError("device_function must be HIP C++ source defining `integrand`, "
                        "an Integrand, or a library integrand name")
    dom = None
    if domain is not None:  # python_wrapper.py:20-46
        start = np.array([float(d[0]) for d in domain])
        rng = np.array([float(d[1]) - float(d[0]) for d in domain])
        if len(start) != n_dim:
            raise ValueError("domain must have n_dim intervals")
        if np.any(start) or not np.all(rng == 1):
            dom = np.concatenate([start, rng])
    if multi_gpu is None:  # python_wrapper.py:63-69
        multi_gpu = n_evals >= 1e10
    if multi_cpu is None:
        multi_cpu = (n_evals // max_it) >= 1e8
    n_eval_it = n_evals // max_it
    integ = Integrator(source, n_dim, n_eval_it, n_intervals=n_intervals, n_strat=n_strat,
                       max_it=max_it, skip=skip, alpha=alpha, beta=beta, seed=seed,
                       precision=precision, run_len=run_len, domain=dom, options=options,
                       cache_program=True)
    try:
        integ.run(max_it, use_graph=use_graph)
        r = integ.result()
        I, S2, ne = integ.history()
    finally:
        integ.close()
    tot_ms = (time.perf_counter() - t0) * 1e3
    info = {
        "error": r["error"], "n_strat": n_strat, "n_cubes": r["n_cubes"],
        "n_intervals": n_intervals, "tot_nevals": r["tot_nevals"], "tot_time": tot_ms,
        "avg_it_time": r["loop_ms"] / max_it, "seed": seed, "multi_gpu": bool(multi_gpu),
        "multi_cpu": bool(multi_cpu), "gpu_num": 1, "omp_threads_num": 0, "gpu_thr": gpu_thr,
        "chi2_dof": r["chi2_dof"], "iterations": list(zip(I.tolist(), np.sqrt(S2).tolist())),
        "error_unbiased": r["error_unbiased"], "chi2_dof_unbiased": r["chi2_dof_unbiased"],
        "evals_per_s": r["tot_nevals"] / (r["loop_ms"] * 1e-3) if r["loop_ms"] > 0 else float("nan"),
        "precision": precision, "compile_ms": r["compile_ms"], "setup_ms": r["setup_ms"],
        "exact": exact,
    }
    return r["result"], info
