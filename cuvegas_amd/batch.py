"""Batches of independent integrals advanced in lockstep, one sampling launch
per iteration for the whole batch (C ABI cvg_batch_*).

The integrand is shared by all items and receives the item's parameters and
dimension:  vg_real integrand(vg_real* x, const vg_real* theta, int d).
Item i draws from Philox stream i, so it reproduces the single-integral run
of the same integrand with stream_id = i.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Sequence

import numpy as np

from ._lib import check, cvg_batch_item, lib
from .api import Program, _dptr, _iptr, _llptr, make_params


class BatchProgram(Program):
    """hipRTC-compiled batched sampling kernel (runtime dimension <= max_dim)."""

    def __init__(self, source: str, max_dim: int, n_intervals: int = 256, precision="fp32",
                 run_len: int = 0, options: Optional[str] = None):
        params, _ = make_params(max_dim, n_intervals, 1, 1, 0, 1, 0.5, 0.75, 0, precision,
                                run_len=run_len)
        h = C.c_void_p()
        rc = lib.cvg_batch_program_create(source.encode(), int(max_dim), C.byref(params),
                                          options.encode() if options else None, C.byref(h))
        self._h = h
        if rc < 0:
            msg = lib.cvg_last_error().decode(errors="replace")
            if h.value:
                lib.cvg_program_destroy(h)
                self._h = C.c_void_p()
            from ._lib import CuvegasError
            raise CuvegasError(rc, msg)
        self.log = lib.cvg_program_log(h).decode(errors="replace")
        self.code_size = lib.cvg_program_code_size(h)
        self.max_dim, self.n_intervals, self.precision, self.run_len = max_dim, n_intervals, precision, run_len


class Batch:
    """items: sequence of (n_dim, n_eval_it, theta_offset[, n_strat])."""

    def __init__(self, program: BatchProgram, items: Sequence, theta: Sequence[float], *,
                 max_it: int = 10, skip: int = 5, alpha: float = 0.5, beta: float = 0.75,
                 seed: int = 0):
        self.program = program
        n = len(items)
        arr = (cvg_batch_item * n)()
        for i, it in enumerate(items):
            arr[i].n_dim = int(it[0])
            arr[i].n_eval_it = int(it[1])
            arr[i].theta_offset = int(it[2])
            arr[i].n_strat = int(it[3]) if len(it) > 3 else 0
        common, _ = make_params(program.max_dim, program.n_intervals, 1, max_it, skip, 1, alpha, beta,
                                seed, program.precision, run_len=program.run_len)
        th = np.ascontiguousarray(theta if len(theta) else [0.0], dtype=np.float64)
        h = C.c_void_p()
        check(lib.cvg_batch_create(program.handle, C.byref(common), arr, n, _dptr(th), len(theta),
                                   C.byref(h)))
        self._h = h
        self.n_items, self.max_it, self.skip = n, max_it, skip
        self.n_intervals = program.n_intervals
        self.dims = [int(it[0]) for it in items]

    def run(self, n_iter: int, use_graph: bool = True) -> "Batch":
        check(lib.cvg_batch_run(self._h, int(n_iter), 1 if use_graph else 0))
        return self

    def results(self) -> dict:
        n = self.n_items
        res, err, chi2 = np.zeros(n), np.zeros(n), np.zeros(n)
        erru, chi2u = np.zeros(n), np.zeros(n)
        ne = np.zeros(n, dtype=np.int64)
        ms = C.c_double()
        check(lib.cvg_batch_results(self._h, _dptr(res), _dptr(err), _dptr(chi2), _llptr(ne), C.byref(ms),
                                    _dptr(erru), _dptr(chi2u)))
        return {"result": res, "error": err, "chi2_dof": chi2, "nevals": ne, "loop_ms": ms.value,
                "error_unbiased": erru, "chi2_dof_unbiased": chi2u}

    def set_timing(self, enable: bool = True) -> None:
        """Time the sampling launch and the whole iteration of eager runs
        (run(..., use_graph=False)) with HIP events on the batch's stream."""
        check(lib.cvg_batch_set_timing(self._h, 1 if enable else 0))

    def timing(self):
        """(fill_ms, iter_ms) of every timed iteration so far."""
        f, t = np.zeros(self.max_it), np.zeros(self.max_it)
        k = check(lib.cvg_batch_get_timing(self._h, _dptr(f), _dptr(t), self.max_it))
        return f[:k], t[:k]

    def history(self, item: int):
        I, S2 = np.zeros(self.max_it), np.zeros(self.max_it)
        ne = np.zeros(self.max_it, dtype=np.int64)
        k = check(lib.cvg_batch_history(self._h, int(item), _dptr(I), _dptr(S2), _llptr(ne), self.max_it))
        return I[:k], S2[:k], ne[:k]

    def item_info(self, item: int):
        d, ns, C_ = C.c_int(), C.c_int(), C.c_longlong()
        check(lib.cvg_batch_item_info(self._h, int(item), C.byref(d), C.byref(ns), C.byref(C_)))
        return d.value, ns.value, C_.value

    def get_map(self, item: int) -> np.ndarray:
        x = np.zeros((self.dims[item], self.n_intervals + 1))
        check(lib.cvg_batch_get_map(self._h, int(item), _dptr(x)))
        return x

    def get_nh(self, item: int) -> np.ndarray:
        nh = np.zeros(self.item_info(item)[2], dtype=np.int32)
        check(lib.cvg_batch_get_nh(self._h, int(item), _iptr(nh)))
        return nh

    def get_cube_sums(self, item: int):
        c = self.item_info(item)[2]
        a, b = np.zeros(c), np.zeros(c)
        check(lib.cvg_batch_get_cube_sums(self._h, int(item), _dptr(a), _dptr(b)))
        return a, b

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.cvg_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --- the serving-style workload: Gaussian peak, oscillatory and corner peak ----

FAMILY_SOURCE = """
__device__ vg_real integrand(vg_real* x, const vg_real* th, int d) {
    const int fam = (int)th[0];
    if (fam == 0) {          // normalised Gaussian peak: sigma, norm, m = -1/(2 sigma^2), centre[d]
        vg_real q = (vg_real)0;
        for (int i = 0; i < d; ++i) { const vg_real a = x[i] - th[4 + i]; q = fma(a, a, q); }
        return th[2] * vg_expr(q * th[3]);
    } else if (fam == 1) {   // oscillatory: u, a[d]
        vg_real t = (vg_real)2 * (vg_real)3.14159265358979323846 * th[1];
        for (int i = 0; i < d; ++i) t = t + th[2 + i] * x[i];
        return cos(t);
    }
    vg_real t = (vg_real)1;  // corner peak: a[d]
    for (int i = 0; i < d; ++i) t = t + th[1 + i] * x[i];
    const vg_real r = (vg_real)1 / t;
    vg_real f = (vg_real)1;
    for (int i = 0; i <= d; ++i) f = f * r;
    return f;
}
"""


def draw_serving_items(n: int = 512, seed: int = 3):
    """Parameter draws of the serving-style workload, as plain dicts: d ~
    U{2..12}; family uniform among Gaussian peak (sigma ~ U(0.02,0.2), centre ~
    U(0.3,0.7)^d), oscillatory (u ~ U(0,1), a_i ~ U(0,5)) and corner peak
    (a_i ~ U(0.5,3)); per-iteration budget log-uniform in [1e4, 1e7]; numpy
    default_rng(seed), in that order per item.  The 512-integral draw of seed 3
    is committed as a fixture (tests/golden/workloads.json, key batch_512) and
    read from there by the benchmark and the tests."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        d = int(rng.integers(2, 13))
        fam = ["gaussian", "oscillatory", "corner"][int(rng.integers(0, 3))]
        neval = int(round(math.exp(rng.uniform(math.log(1e4), math.log(1e7)))))
        if fam == "gaussian":
            p = {"sigma": float(rng.uniform(0.02, 0.2)), "centre": rng.uniform(0.3, 0.7, d).tolist()}
        elif fam == "oscillatory":
            p = {"u": float(rng.uniform(0, 1)), "a": rng.uniform(0, 5, d).tolist()}
        else:
            p = {"a": rng.uniform(0.5, 3.0, d).tolist()}
        out.append({"d": d, "family": fam, "neval": neval, **p})
    return out


def workload_from_items(entries):
    """Batch items, the parameter block of FAMILY_SOURCE and the closed forms
    for a list of item dicts ({"d", "family", "neval"} plus the family's
    parameters, as draw_serving_items returns and the fixture holds).
    Returns (items, theta, exact, family names); item i's parameters start at
    theta[items[i][2]]: [0, sigma, norm, -1/(2 sigma^2), centre[d]],
    [1, u, a[d]] or [2, a[d]]."""
    from . import integrands as I
    items, theta, exact, fams = [], [], [], []
    for e in entries:
        d, fam, neval = int(e["d"]), e["family"], int(e["neval"])
        off = len(theta)
        if fam == "gaussian":
            sigma, centre = float(e["sigma"]), [float(c) for c in e["centre"]]
            norm = (1.0 / (sigma * math.sqrt(2.0 * math.pi))) ** d
            theta += [0.0, sigma, norm, -1.0 / (2.0 * sigma * sigma), *centre]
            exact.append(math.prod(I._erf_mass(c, sigma) for c in centre))
        elif fam == "oscillatory":
            u, a = float(e["u"]), [float(v) for v in e["a"]]
            theta += [1.0, u, *a]
            exact.append(I.oscillatory(d, u, a).exact)
        elif fam == "corner":
            a = [float(v) for v in e["a"]]
            theta += [2.0, *a]
            exact.append(I.corner_peak(d, a).exact)
        else:
            raise ValueError(f"unknown family {fam!r}")
        items.append((d, neval, off))
        fams.append(fam)
    return items, theta, np.array(exact), fams


def serving_workload(n: int = 512, seed: int = 3):
    """draw_serving_items(n, seed) as (items, theta, exact, family names)."""
    return workload_from_items(draw_serving_items(n, seed))
