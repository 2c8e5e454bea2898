"""Backward replay of parametrised integrands: against the oracle's replay of
the same frozen iteration, central finite differences of the frozen estimator,
and the closed-form erf gradient of the Gaussian mixture."""
import json
import math
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print("[gpu-case]", *a, flush=True)


def _theta():
    with open(os.path.join(GOLD, "workloads.json")) as f:
        return np.array(json.load(f)["mixture_6d"]["theta"])


def _state(prec, theta, N=400_000, iters=8, skip=4, ns=None, nbins=256, seed=2):
    from cuvegas_amd import Integrator
    from cuvegas_amd import integrands as I
    itg = I.gaussian_mixture_param(6, theta)
    st = Integrator(itg.source, 6, N, n_intervals=nbins, n_strat=ns, max_it=iters, skip=skip,
                    seed=seed, precision=prec, theta=theta, options="grad")
    st.set_record(True)
    return st, itg


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_replay_matches_oracle(orc, prec):
    """One recorded iteration: value, variance, gradient and gradient variance
    against the oracle's replay of the same frozen map and allocation."""
    th = _theta()
    _say("grad-oracle", prec)
    st, itg = _state(prec, th, N=60_000, iters=3, skip=2, ns=5, nbins=64)
    st.run(3)
    v, vv, g, gv = st.backward()
    x, nh = st.snapshot(2)
    p, keep = orc.make_params(6, 64, 5, 60_000, 3, orc.F_MIXTURE, list(th), skip=2, seed=2,
                              precision=0 if prec == "fp64" else 1)
    oI, oS2, og, ogv = orc.grad_replay(p, keep, x, nh, 2, th)
    tol = 1e-11 if prec == "fp64" else 2e-5
    assert abs(v - oI) <= tol * abs(oI)
    assert abs(vv - oS2) <= 1e-6 * abs(oS2) + 1e-300
    assert np.max(np.abs(g - og) / (np.abs(og) + 1e-12 * np.abs(og).max())) < tol * 10
    assert np.max(np.abs(gv - ogv) / (np.abs(ogv) + 1e-30)) < 1e-4
    # the replay at the forward's theta reproduces the forward estimate
    I_k, S2, ne = st.history()
    assert abs(v - I_k[2]) <= 1e-12 * abs(I_k[2]) + (1e-6 * abs(I_k[2]) if prec == "fp32" else 0)


@pytest.mark.parametrize("d", [2, 9])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_replay_matches_oracle_other_dims(orc, prec, d):
    """The replay for other parameter counts (2-D: 16 parameters; 9-D: 44,
    more than one chunk of the kernel's transposed wave sums) against the
    oracle's replay of the same frozen iteration."""
    from cuvegas_amd import Integrator
    from cuvegas_amd import integrands as I
    rng = np.random.default_rng(10 + d)
    th = np.concatenate([rng.uniform(0.5, 1.5, 4), rng.uniform(0.05, 0.15, 4), rng.uniform(0.2, 0.8, 4 * d)])
    _say("grad-oracle-dims", prec, d)
    ns, nb, N = (40, 64, 40_000) if d == 2 else (2, 32, 60_000)
    st = Integrator(I.gaussian_mixture_param(d, th).source, d, N, n_intervals=nb, n_strat=ns, max_it=3,
                    skip=2, seed=5, precision=prec, theta=th, options="grad")
    st.set_record(True)
    st.run(3)
    v, vv, g, gv = st.backward()
    x, nh = st.snapshot(2)
    p, keep = orc.make_params(d, nb, ns, N, 3, orc.F_MIXTURE, list(th), skip=2, seed=5,
                              precision=0 if prec == "fp64" else 1)
    oI, oS2, og, ogv = orc.grad_replay(p, keep, x, nh, 2, th)
    tol = 1e-11 if prec == "fp64" else 2e-5
    assert abs(v - oI) <= tol * abs(oI)
    assert abs(vv - oS2) <= 1e-6 * abs(oS2) + 1e-300
    assert np.max(np.abs(g - og) / (np.abs(og) + 1e-12 * np.abs(og).max())) < tol * 10
    assert np.max(np.abs(gv - ogv) / (np.abs(ogv) + 1e-30)) < 1e-4


def test_gradient_vs_finite_differences():
    """d value / d theta of the frozen estimator vs central differences (h = 1e-4)."""
    th = _theta()
    _say("grad-fd")
    st, itg = _state("fp64", th, N=200_000, iters=6, skip=3)
    st.run(6)
    v0, _, g, _ = st.backward()
    def fd(p, h):
        tp, tm = th.copy(), th.copy()
        tp[p] += h
        tm[p] -= h
        return (st.backward(tp)[0] - st.backward(tm)[0]) / (2 * h)

    scale = np.abs(g).max()
    for p in range(th.size):
        f1 = fd(p, 1e-4)  # central difference: truncation ~(h / s_k)^2, up to ~1e-5 for s_k = 0.037
        assert abs(f1 - g[p]) <= 2e-4 * max(abs(g[p]), 1e-3 * scale), (p, f1, g[p])
        # Richardson extrapolation removes the h^2 term: agreement to ~1e-8
        f2 = fd(p, 5e-5)
        rich = (4 * f2 - f1) / 3
        assert abs(rich - g[p]) <= 1e-7 * max(abs(g[p]), 1e-3 * scale), (p, rich, g[p])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gradient_over_domain_vs_finite_differences(prec):
    """Replay over a sub-box (x = start + range * y, values and gradients
    scaled by the box volume): the value at the forward's theta equals the
    forward's weighted mean of the recorded iterations, and the gradient
    matches central differences of the frozen estimator."""
    from cuvegas_amd import Integrator
    from cuvegas_amd import integrands as I
    th = _theta()
    _say("grad-domain", prec)
    dom = [[0.1, 0.15, 0.1, 0.2, 0.1, 0.05], [0.8, 0.7, 0.85, 0.7, 0.8, 0.9]]
    st = Integrator(I.gaussian_mixture_param(6, th).source, 6, 200_000, n_intervals=256, max_it=6, skip=3,
                    seed=7, precision=prec, theta=th, options="grad", domain=dom)
    st.set_record(True)
    st.run(6)
    v0, _, g, _ = st.backward()
    I_k, S2, _ = st.history()
    w = 1.0 / np.asarray(S2[3:6])
    mean = float((w * np.asarray(I_k[3:6])).sum() / w.sum())
    assert abs(v0 - mean) <= (1e-12 if prec == "fp64" else 1e-6) * abs(mean)
    scale = np.abs(g).max()
    # FP32: the frozen estimator (|v| ~ 3.4) carries ~1e-7 relative rounding
    # noise, ~1e-4 after dividing by 2h = 4e-3, hence the absolute floor
    h = 1e-4 if prec == "fp64" else 2e-3
    tol = 2e-4 if prec == "fp64" else 2e-2
    floor = 2e-7 * scale if prec == "fp64" else 1e-4 * scale
    for p in (0, 3, 5, 8, 13, 27):
        tp, tm = th.copy(), th.copy()
        tp[p] += h
        tm[p] -= h
        fd = (st.backward(tp)[0] - st.backward(tm)[0]) / (2 * h)
        assert abs(fd - g[p]) <= tol * abs(g[p]) + floor, (p, fd, g[p])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gradient_vs_closed_form(prec):
    """The 6-D mixture configuration (5e7 evaluations over 10 iterations: 5e6
    per iteration, 1024 bins, n_strat 10 by the reference's formula, 5
    iterations averaged and replayed): every one of the 32 gradient components
    within 3 sigma of the closed-form erf gradient, the squared pulls
    averaging below 2, the value within 4 sigma of its closed form."""
    th = _theta()
    _say("grad-closed-form", prec)
    st, itg = _state(prec, th, N=5_000_000, iters=10, skip=5, nbins=1024, seed=0)
    st.run(10)
    r = st.result()
    v, vv, g, gv = st.backward()
    exact_v = itg.exact
    exact_g = np.array(itg.exact_grad(th))
    assert abs(r["result"] - exact_v) < 4 * r["error"]
    assert abs(v - exact_v) < 4 * math.sqrt(vv)
    pulls = (g - exact_g) / np.sqrt(gv)
    print("max |pull|", np.abs(pulls).max(), "within 3 sigma", np.mean(np.abs(pulls) < 3))
    assert st.n_strat == 10 and st.n_cubes == 10 ** 6
    assert np.all(np.abs(pulls) < 3), pulls
    assert np.mean(pulls ** 2) < 2.0


def test_torch_autograd_function():
    import torch
    from cuvegas_amd import integrands as I
    from cuvegas_amd.autograd import DifferentiableIntegral
    th = _theta()
    _say("torch-autograd")
    itg = I.gaussian_mixture_param(6, th)
    mod = DifferentiableIntegral(itg.source, 6, th.size, 300_000, max_it=6, skip=3, n_intervals=256,
                                 seed=4)
    theta = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    val = mod(theta)
    val.backward()
    v, _, g, _ = mod.state.backward()
    assert val.item() == pytest.approx(mod.last["result"], rel=0, abs=0)
    assert np.allclose(theta.grad.numpy(), g, rtol=0, atol=0)
    assert abs(val.item() - itg.exact) < 4 * mod.last["error"]
    # chain rule through a downstream op
    theta.grad = None
    (2.0 * mod(theta)).backward()
    assert np.allclose(theta.grad.numpy(), 2.0 * mod.state.backward()[2])


def test_forward_releases_superseded_state():
    """A new forward closes the previous forward's integrator once no pending
    backward needs it; an integrator whose backward is still pending stays
    usable (loss = F(t1) + F(t2) differentiates both)."""
    import torch
    from cuvegas_amd import integrands as I
    from cuvegas_amd.autograd import DifferentiableIntegral
    th = _theta()
    _say("torch-autograd-release")
    itg = I.gaussian_mixture_param(6, th)
    mod = DifferentiableIntegral(itg.source, 6, th.size, 100_000, max_it=4, skip=2, n_intervals=64, seed=4)
    t1 = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    t2 = torch.tensor(th * 1.01, dtype=torch.float64, requires_grad=True)
    loss = mod(t1) + mod(t2)
    first = next(iter(mod._pending - {mod.state}))
    assert first._h.value  # pending: not closed
    loss.backward()
    assert t1.grad is not None and t2.grad is not None and torch.all(torch.isfinite(t1.grad))
    kept = mod.state
    (mod(t1)).backward()
    assert not kept._h.value  # superseded, no backward pending: released at the new forward
    assert mod.state._h.value


def test_parametrised_state_save_restore(tmp_path):
    """A parametrised integrator's archive keeps theta: the restored integrator
    continues with the saved parameters (same next iteration as the original),
    and an integrator with another parameter count refuses the archive."""
    from cuvegas_amd import Integrator
    from cuvegas_amd import integrands as I
    th = _theta()
    _say("grad-save-restore")
    itg = I.gaussian_mixture_param(6, th)
    a = Integrator(itg.source, 6, 100_000, n_intervals=64, max_it=5, skip=2, seed=3, theta=th)
    a.run(3)
    path = str(tmp_path / "mix.npz")
    a.save(path)
    a.run(1)
    b = Integrator.from_state(path, itg.source)
    assert np.array_equal(b.theta, th)
    b.run(1)
    assert np.array_equal(a.history()[0], b.history()[0]) and np.array_equal(a.get_nh(), b.get_nh())
    with pytest.raises(ValueError):
        Integrator(I.gaussian_peak(6).source, 6, 100_000, n_intervals=64, max_it=5, skip=2).restore(path)
    a.close()
    b.close()
