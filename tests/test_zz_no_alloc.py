"""cvg_state_backward_noalloc, the backward replay that allocates nothing,
through the C ABI: bit for bit the results of cvg_state_backward on two calls
in a row (the second on a smaller state, so the shared static scratch still
holds the first call's contents), at another theta, and the refusals (a
replay that needs more than the static scratch; a program without the
gradient kernel)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)


def _mixture_state(d, n_eval_it, nbins, n_strat=None, seed=3, iters=6, skip=3, prec="fp64"):
    from cuvegas_amd import Integrator
    from cuvegas_amd import integrands as I
    rng = np.random.default_rng(50 + d)
    th = np.concatenate([rng.uniform(0.5, 1.5, 4), rng.uniform(0.05, 0.15, 4), rng.uniform(0.2, 0.8, 4 * d)])
    st = Integrator(I.gaussian_mixture_param(d, th).source, d, n_eval_it, n_intervals=nbins, n_strat=n_strat,
                    max_it=iters, skip=skip, seed=seed, precision=prec, theta=th, options="grad")
    st.set_record(True)
    st.run(iters)
    return st, th


def _backward(fn, st, theta=None):
    from cuvegas_amd._lib import check
    P = st.params.n_theta
    g, gv = np.zeros(P), np.zeros(P)
    v, vv = C.c_double(), C.c_double()
    th = None if theta is None else np.ascontiguousarray(theta, dtype=np.float64)
    dp = C.POINTER(C.c_double)
    check(fn(st._h, None if th is None else th.ctypes.data_as(dp), g.ctypes.data_as(dp), gv.ctypes.data_as(dp),
             C.byref(v), C.byref(vv)))
    return v.value, vv.value, g, gv


def _same(a, b):
    return a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_backward_noalloc_matches_backward(prec):
    from cuvegas_amd._lib import lib
    _say("backward-noalloc", prec)
    big, th_big = _mixture_state(6, 300_000, 256, prec=prec)
    small, th_small = _mixture_state(3, 40_000, 64, prec=prec)
    assert big.n_cubes > small.n_cubes
    ref_big = _backward(lib.cvg_state_backward, big)
    ref_small = _backward(lib.cvg_state_backward, small)
    # first call on this device, then a smaller state over the same scratch
    assert _same(_backward(lib.cvg_state_backward_noalloc, big), ref_big)
    assert _same(_backward(lib.cvg_state_backward_noalloc, small), ref_small)
    # the replay at another theta (parameters staged in the scratch as well)
    th2 = th_big * 1.001
    assert _same(_backward(lib.cvg_state_backward_noalloc, big, th2), _backward(lib.cvg_state_backward, big, th2))
    assert _same(_backward(lib.cvg_state_backward_noalloc, small), ref_small)
    big.close()
    small.close()


def test_backward_noalloc_refusals():
    from cuvegas_amd import Integrator
    from cuvegas_amd import integrands as I
    from cuvegas_amd._lib import CVG_ERR_CAPACITY, CVG_ERR_INVALID, lib
    _say("backward-noalloc-refusals")
    # 14^6 = 7.5 M cubes: the replay's prefix sums alone take 60 MB, with the
    # unit partials more than the 64 MiB of static scratch
    st, _ = _mixture_state(6, 2 * 14 ** 6, 64, n_strat=14, iters=2, skip=1)
    dp = C.POINTER(C.c_double)
    g, gv, v, vv = np.zeros(32), np.zeros(32), C.c_double(), C.c_double()
    rc = lib.cvg_state_backward_noalloc(st._h, None, g.ctypes.data_as(dp), gv.ctypes.data_as(dp), C.byref(v),
                                        C.byref(vv))
    assert rc == CVG_ERR_CAPACITY
    assert b"64" in lib.cvg_last_error() or b"67108864" in lib.cvg_last_error()
    st.close()
    # a program without the gradient kernel
    plain = Integrator(I.gaussian_peak(3).source, 3, 20_000, n_intervals=32, max_it=2, skip=1)
    plain.run(2)
    rc = lib.cvg_state_backward_noalloc(plain._h, None, None, None, None, None)
    assert rc == CVG_ERR_INVALID
    plain.close()
