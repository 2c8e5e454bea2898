"""The CPU oracle pinned against published vectors, libm, exact arithmetic and
closed forms (no GPU needed)."""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_philox_known_answers(orc):
    with open(os.path.join(GOLD, "philox4x32_10_kat.json")) as f:
        kat = json.load(f)["vectors"]
    for v in kat:
        ctr = [int(x, 16) for x in v["ctr"]]
        key = [int(x, 16) for x in v["key"]]
        out = orc.philox(ctr, key)
        assert [f"{w:08x}" for w in out] == v["out"]


def test_log_exp_pow_against_libm(orc):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(1e-3, 10, 2000), 10.0 ** rng.uniform(-300, 300, 2000),
                         [1e-310, 5e-324, 2.2250738585072014e-308, 0.5, 1.0, 2.0, 0.75,
                          1.4142135623730951, 0.7071067811865476]])
    for x in xs:
        ref = math.log(x)
        got = orc.log(x)
        assert abs(got - ref) <= 4e-16 * max(abs(ref), 1e-300) + 1e-320, (x, got, ref)
    ys = np.concatenate([rng.uniform(-700, 700, 3000), rng.uniform(-1, 1, 1000),
                         [0.0, -745.0, 709.7, 1e-20, -1e-20]])
    for y in ys:
        ref = math.exp(y)
        got = orc.exp(y)
        assert abs(got - ref) <= 4e-16 * ref + 5e-324, (y, got, ref)
    for _ in range(3000):
        x = 10.0 ** rng.uniform(-250, 5)
        a = rng.uniform(0.01, 2.0)
        ref = x ** a
        assert abs(orc.pow_(x, a) - ref) <= 1e-13 * ref + 1e-320, (x, a)
    # special values
    assert orc.pow_(0.0, 0.75) == 0.0
    assert orc.pow_(1.0, 0.3) == 1.0
    assert orc.pow_(7.0, 0.0) == 1.0
    assert math.isnan(orc.pow_(-1.0, 0.5))
    assert orc.log(1.0) == 0.0 and orc.exp(0.0) == 1.0
    assert orc.log(0.0) == -math.inf and math.isnan(orc.log(-1.0))


def test_sat_int(orc):
    assert orc.sat_int(float("nan")) == 0
    assert orc.sat_int(3.99) == 3 and orc.sat_int(-3.99) == -3
    assert orc.sat_int(1e300) == 2**31 - 1 and orc.sat_int(-1e300) == -(2**31)


@pytest.mark.parametrize("nbins,nstrat", [(1024, 8), (1024, 22), (512, 2), (4, 3), (7, 5), (1000, 1)])
def test_bin_of_matches_exact_rational_floor(orc, nbins, nstrat):
    """id = floor(y*n) and offset = y*n - id for y = (k + (2r+1)/2^33) / nstrat, exactly."""
    rng = np.random.default_rng(nbins * 31 + nstrat)
    base = list(rng.integers(0, 2**32, 300)) + [0, 1, 2**31 - 1, 2**31, 2**32 - 1, 2**32 - 2]
    for k in sorted({0, nstrat - 1, nstrat // 2}):
        # words straddling every bin boundary inside stratum k: u = j*nstrat/nbins - k
        words = list(base)
        for j in range(k * nbins // nstrat, (k + 1) * nbins // nstrat + 2):
            u = Fraction(j * nstrat, nbins) - k
            if 0 < u < 1:
                w = math.floor(u * 2**32 - Fraction(1, 2))
                words += [v for v in (w - 1, w, w + 1) if 0 <= v < 2**32]
        for r in words[:2000]:
            r = int(r)
            y = (Fraction(k) + Fraction(2 * r + 1, 2**33)) / nstrat
            yn = y * nbins
            exp_id = math.floor(yn)
            i, rem, low = orc.bin_of(r, k * nbins, nbins, nstrat)
            assert i == exp_id
            assert Fraction(rem) + Fraction(low, 2**33) == (yn - exp_id) * nstrat
            assert 0 <= i < nbins


def test_offset_single_conversion_rewrite_is_exact():
    """The sampling kernel forms the in-bin offset with one integer-to-float
    conversion when n_strat <= 256:
      FP32  (float)(rem << 24 | low >> 9) * (2^-24 / ns)
            == ((float)rem + (float)(low >> 9) * 2^-24) * (1 / ns)
      FP64  fma((double)(rem << 1 | low >> 32), 2^32, (double)(low mod 2^32)) * (2^-33 / ns)
            == ((double)rem + (double)low * 2^-33) * (1 / ns)
    bit for bit, for every ns in 2..256, every rem < ns and any 33-bit low."""
    rng = np.random.default_rng(5)
    f32, f64 = np.float32, np.float64
    edge = np.array([0, 1, 2**9 - 1, 2**9, 2**32 - 1, 2**32, 2**33 - 1, 2**33 - 2**9], dtype=np.uint64)
    for ns in range(2, 257):
        rem = np.concatenate([rng.integers(0, ns, 400), np.arange(ns), np.full(8, ns - 1)]).astype(np.uint64)
        low = np.concatenate([rng.integers(0, 2**33, 400), rng.integers(0, 2**33, ns), edge]).astype(np.uint64)
        inv32 = f32(1) / f32(ns)
        ref32 = (rem.astype(f32) + (low >> np.uint64(9)).astype(f32) * f32(2.0**-24)) * inv32
        one32 = ((rem << np.uint64(24)) | (low >> np.uint64(9))).astype(np.uint32).astype(f32) * (f32(2.0**-24) * inv32)
        assert np.array_equal(ref32.view(np.uint32), one32.view(np.uint32)), ns
        inv64 = 1.0 / ns
        ref64 = (rem.astype(f64) + low.astype(f64) * 2.0**-33) * inv64
        hi = ((rem << np.uint64(1)) | (low >> np.uint64(32))).astype(f64)
        v = hi * 2.0**32 + (low & np.uint64(0xFFFFFFFF)).astype(f64)  # exact: < 2^41
        one64 = v * (2.0**-33 * inv64)
        assert np.array_equal(ref64.view(np.uint64), one64.view(np.uint64)), ns


def test_scaled_table_form_is_exact():
    """For a power-of-two bin count n the sampling tables hold (edge, n * width)
    and the in-bin offset carries the 1/n: the kernel's offset constant
    (2^-24 / ns) * (1/n) and the table width n * width are exact rescalings, so
    (n * width) * (offset / n) is the same real number as width * offset (the
    point's fma rounds it once, identically) and the Jacobian factor n * width
    is the value the unscaled form multiplies out.  Checked over FP32 and FP64
    offsets, widths down to the smallest subnormals, and n = 2 .. 2^14 (other
    bin counts keep the unscaled form)."""
    rng = np.random.default_rng(11)
    f32, f64 = np.float32, np.float64
    ns = 8
    for n in [2 ** k for k in range(1, 15)]:
        # FP32: offset = bits * c with c = 2^-24 / ns, scaled by 1/n in the constant
        bits = rng.integers(0, 2**32, 4000, dtype=np.uint64).astype(np.uint32)
        c = f32(2.0**-24) * (f32(1) / f32(ns))
        off = bits.astype(f32) * c
        off_s = bits.astype(f32) * (c * (f32(1) / f32(n)))
        assert np.array_equal((off_s * f32(n)).view(np.uint32), off.view(np.uint32)), n
        w = np.concatenate([rng.random(2000).astype(f32) * f32(2.0**-20),
                            np.array([np.finfo(f32).smallest_subnormal, f32(1e-40), f32(1.0)], dtype=f32)])
        w_s = w * f32(n)
        assert np.array_equal((w_s / f32(n)).view(np.uint32), w.view(np.uint32)), n
        # products of the two forms agree exactly (both exact in FP64)
        m = min(len(w), len(off))
        assert np.array_equal(w_s[:m].astype(f64) * off_s[:m].astype(f64),
                              w[:m].astype(f64) * off[:m].astype(f64)), n
        # FP64: offset constant 2^-33 / ns scaled by 1/n; widths scaled by n
        v = rng.integers(0, 2**41, 4000).astype(f64)
        c64 = 2.0**-33 * (1.0 / ns)
        assert np.array_equal((v * (c64 * (1.0 / n))) * n, v * c64), n
        w64 = np.concatenate([rng.random(2000) * 2.0**-40, [5e-324, 1e-310, 1.0]])
        assert np.array_equal((w64 * n) / n, w64), n


def test_tree_sum_exact_and_close(orc):
    rng = np.random.default_rng(3)
    for m in [1, 5, 1023, 1024, 1025, 4096 + 17, 1024 * 1024 + 3]:
        ints = rng.integers(-1000, 1000, m).astype(np.float64)
        assert orc.tree_sum(ints) == float(ints.sum())
    a = rng.standard_normal(200000) * 10.0 ** rng.uniform(-5, 5, 200000)
    assert abs(orc.tree_sum(a) - math.fsum(a)) <= 1e-13 * np.abs(a).sum()
    assert orc.tree_sum(np.zeros(0)) == 0.0


@pytest.mark.parametrize("K", [1, 2, 3, 16, 64])
def test_cube_sums_partition(orc, K):
    """Every sample lands in exactly its cube: exact on integer data, for
    layouts with cubes inside a run, across runs and across windows."""
    rng = np.random.default_rng(K)
    for trial in range(6):
        C = int(rng.integers(1, 400))
        if trial % 3 == 0:
            nh = rng.integers(2, 6, C)
        elif trial % 3 == 1:
            nh = rng.integers(2, 3000, C)
        else:
            nh = np.where(rng.random(C) < 0.1, rng.integers(2000, 20000, C), 2)
        S = int(nh.sum())
        v1 = rng.integers(-50, 50, S).astype(np.float64)
        v2 = v1 * v1
        JF, JF2 = orc.cube_sums(nh, v1, v2, K)
        starts = np.concatenate([[0], np.cumsum(nh)[:-1]])
        ref1 = np.add.reduceat(v1, starts)
        ref2 = np.add.reduceat(v2, starts)
        assert np.array_equal(JF, ref1)
        assert np.array_equal(JF2, ref2)
        # floating data: within the error bound of any summation order
        f1 = rng.standard_normal(S)
        JFf, _ = orc.cube_sums(nh, f1, f1 * f1, K)
        for h in rng.integers(0, C, 20):
            seg = f1[starts[h]:starts[h] + nh[h]]
            assert abs(JFf[h] - math.fsum(seg)) <= 4 * nh[h] * 1.2e-16 * np.abs(seg).sum() + 1e-300


def test_results_constant_integrand(orc):
    """Constant J f = c in a cube: I_h = V c, sigma_h^2 = 0, d_h = 0."""
    C = 7
    nh = np.full(C, 5, dtype=np.int32)
    c = 3.0
    JF = np.full(C, 5 * c)
    JF2 = np.full(C, 5 * c * c)
    V = 1.0 / 8
    dh, sums = orc.results(V, 0.75, nh, JF, JF2)
    assert np.all(dh == 0.0)
    assert sums[0] == pytest.approx(C * V * c, rel=1e-15)
    assert abs(sums[1]) < 1e-30
    # two-point cube: values a, b -> sigma^2 = V^2 (a-b)^2 / 4
    a, b = 1.0, 3.0
    dh, sums = orc.results(0.5, 1.0, np.array([2], dtype=np.int32), np.array([a + b]),
                           np.array([a * a + b * b]))
    assert dh[0] == pytest.approx(0.25 * (a - b) ** 2 / 4, rel=1e-15)
    assert sums[1] == pytest.approx(0.25 * (a - b) ** 2 / 4 / 2, rel=1e-15)
    assert sums[3] == pytest.approx(0.25 * (a - b) ** 2 / 4, rel=1e-15)  # / (n_h - 1) = / 1


def test_results_unbiased_variance_is_the_sample_variance_of_the_mean(orc):
    """sums[3] = sum_h sigma_h^2 / (n_h - 1) is, cube by cube, the unbiased
    estimate of the variance of I_h = V mean(f): the sample variance
    (statistics.variance, divisor n - 1) of the V f values over n_h; sums[1]
    (the reference's) is that times (n_h - 1) / n_h."""
    import statistics
    rng = np.random.default_rng(4)
    nh = np.array([2, 3, 5, 17, 2, 9], dtype=np.int32)
    V = 0.125
    vals = [rng.normal(1.0, 0.7, n) for n in nh]
    JF = np.array([v.sum() for v in vals])
    JF2 = np.array([(v * v).sum() for v in vals])
    _, sums = orc.results(V, 0.75, nh, JF, JF2)
    unb = sum(statistics.variance((V * v).tolist()) / n for v, n in zip(vals, nh))
    ref = sum(statistics.pvariance((V * v).tolist()) / n for v, n in zip(vals, nh))
    assert sums[3] == pytest.approx(unb, rel=1e-12)
    assert sums[1] == pytest.approx(ref, rel=1e-12)


def test_alloc_invariants(orc):
    rng = np.random.default_rng(5)
    C, N = 1000, 100000
    dh = rng.random(C) ** 4
    D = orc.tree_sum(dh)
    nh = orc.alloc(dh, D, N)
    assert nh.min() >= 2
    assert nh.sum() <= N + 2 * C
    expect = np.maximum(2, np.floor(dh / D * N)).astype(np.int64)
    assert np.abs(nh - expect).max() <= 1  # floor of a correctly rounded quotient
    uni = orc.alloc(np.ones(C), float(C), N)
    assert np.all(uni == N // C)
    one_hot = np.zeros(C)
    one_hot[17] = 2.5
    oh = orc.alloc(one_hot, 2.5, N)
    assert oh[17] == N and np.all(np.delete(oh, 17) == 2)
    # D = 0 (constant integrand): 0/0 -> NaN -> (int) 0 -> 2
    assert np.all(orc.alloc(np.zeros(C), 0.0, N) == 2)


def _inverse_cdf_edges(alpha, w, counts, x):
    """Independent construction: damped smoothed weights define a piecewise
    constant density on the old bins; new edges split its mass in n equal
    parts (numpy interpolation of the cumulative mass)."""
    n = w.size
    ww = np.where(counts > 0, w / np.maximum(counts, 1), w)
    s = ww.sum()
    sm = np.empty(n)
    sm[0] = (7 * ww[0] + ww[1]) / (8 * s)
    sm[-1] = (ww[-2] + 7 * ww[-1]) / (8 * s)
    sm[1:-1] = (ww[:-2] + 6 * ww[1:-1] + ww[2:]) / (8 * s)
    d = np.where(sm != 0, ((sm - 1) / np.log(np.where(sm != 0, sm, 0.5))) ** alpha, 0.0)
    cum = np.concatenate([[0.0], np.cumsum(d)])
    targets = np.arange(1, n) * cum[-1] / n
    return np.concatenate([[x[0]], np.interp(targets, cum, x), [x[-1]]])


def test_refine_fixed_point_and_inverse_cdf(orc):
    n = 64
    x = np.linspace(0.0, 1.0, n + 1)
    # uniform weights leave the map unchanged
    out = orc.refine_dim(0.5, np.ones(n), np.ones(n, dtype=np.uint32), x)
    assert np.max(np.abs(out - x)) <= 1e-15
    assert out[0] == 0.0 and out[-1] == 1.0
    # alpha = 0 damps every non-zero weight to 1: uniform bins again
    rng = np.random.default_rng(2)
    w = rng.random(n) + 0.1
    out0 = orc.refine_dim(0.0, w, np.ones(n, dtype=np.uint32), x)
    assert np.max(np.abs(out0 - x)) <= 1e-14
    # general weights on a non-uniform map: equal-mass inverse CDF
    xr = np.concatenate([[0.0], np.sort(rng.random(n - 1)), [1.0]])
    for alpha in (0.5, 1.0, 1.5):
        w = rng.random(n) ** 3 * 10
        c = rng.integers(1, 50, n).astype(np.uint32)
        out = orc.refine_dim(alpha, w * c, c, xr)
        ref = _inverse_cdf_edges(alpha, w * c, c, xr)
        assert np.max(np.abs(out - ref)) <= 1e-12
        assert np.all(np.diff(out) > 0)
    # zero weights: map unchanged (the reference would produce NaN)
    assert np.array_equal(orc.refine_dim(0.5, np.zeros(n), np.zeros(n, dtype=np.uint32), xr), xr)


def test_run_closed_form_gaussian(orc):
    """4-D normalised Gaussian (sigma = 0.1): pull within 4 sigma, chi2/dof < 3."""
    d, n, ns, N, iters = 4, 256, 6, 20000, 10
    sigma, mu = 0.1, 0.5
    norm = (1.0 / (sigma * math.sqrt(2 * math.pi))) ** d
    m = -1.0 / (2.0 * sigma * sigma)
    p, keep = orc.make_params(d, n, ns, N, iters, orc.F_GAUSS, (mu, sigma, norm, m), seed=0)
    r = orc.run(p, keep, ns ** d)
    exact = (math.erf(0.5 / (sigma * math.sqrt(2)))) ** d
    assert abs(r["result"] - exact) < 4 * r["error"]
    assert r["chi2_dof"] < 3.0
    assert r["error"] / exact < 2e-3
    assert r["tot_nevals"] == r["hist_n"].sum()
    assert np.all(r["nh"] >= 2)
    assert np.all(np.diff(r["x_edges"], axis=1) > 0)


def test_run_polynomial_fp32_and_domain(orc):
    d, n, ns, N, iters = 3, 64, 5, 5000, 8
    for prec in (0, 1):
        p, keep = orc.make_params(d, n, ns, N, iters, orc.F_POLY, precision=prec, seed=7)
        r = orc.run(p, keep, ns ** d)
        assert abs(r["result"] - 1.0) < 4 * r["error"]
    # linear over [1,3] x [0,2] x [-1,0]: integral = vol * mean
    dom = np.array([1.0, 0.0, -1.0, 2.0, 2.0, 1.0])
    p, keep = orc.make_params(d, n, ns, N, iters, orc.F_LINEAR, seed=3, domain=dom)
    r = orc.run(p, keep, ns ** d)
    exact = 4.0 * (2.0 + 1.0 - 0.5)
    assert abs(r["result"] - exact) < 4 * r["error"] + 1e-12


def test_run_constant_integrand_is_exact(orc):
    """f = 1 in the FP32 form: J f is exactly 1 under the uniform map, so the
    first iterations have zero variance (later ones move the map: with D = 0
    every cube gets the minimum 2 samples and some bins stay empty).  With an
    exact iteration among those averaged, the combine returns the exact
    iterations' mean with error 0 (the reference's inverse-variance sums would
    give inf / inf = NaN); without one, the usual weighted mean."""
    d, n, ns, N, iters = 3, 32, 4, 4000, 6
    p, keep = orc.make_params(d, n, ns, N, iters, orc.F_ONE, precision=1, skip=2)
    r = orc.run(p, keep, ns ** d)
    S2 = r["hist_S2"]
    assert S2[0] == 0.0 and S2[2] == 0.0 and np.all(r["hist_I"][S2 == 0.0] == 1.0)
    assert r["result"] == 1.0 and r["error"] == 0.0 and r["chi2_dof"] == 0.0
    if np.all(S2[4:] > 0):
        p, keep = orc.make_params(d, n, ns, N, iters, orc.F_ONE, precision=1, skip=4)
        r = orc.run(p, keep, ns ** d)
        w = 1.0 / S2[4:]
        assert r["result"] == pytest.approx(np.sum(r["hist_I"][4:] * w) / np.sum(w), rel=1e-14)
        assert np.isfinite(r["error"]) and r["error"] > 0


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_run_closed_form_oscillatory(orc, prec, seed):
    """5-D cos(2 pi u + sum a_i x_i), a_i ~ U(0,5): the oracle integrator (both
    forms) within 4 sigma of cos(2 pi u + sum a_i / 2) prod sin(a_i/2)/(a_i/2),
    chi2/dof < 3.  A dropped phase term or a wrong factor 1/2 inside the
    product moves the result by far more than 4 sigma."""
    from cuvegas_amd import integrands as I
    d, n, ns, N, iters = 5, 128, 4, 40000, 10
    rng = np.random.default_rng(20 + seed)
    itg = I.oscillatory(d, float(rng.uniform(0, 1)), rng.uniform(0, 5, d).tolist())
    p, keep = orc.make_params(d, n, ns, N, iters, orc.F_OSC, list(itg.params), seed=seed, precision=prec)
    r = orc.run(p, keep, ns ** d, snapshots=False)
    assert abs(r["result"] - itg.exact) < 4 * r["error"], (r["result"], itg.exact, r["error"])
    assert r["chi2_dof"] < 3.0
    assert r["error"] < 2e-3  # |f| <= 1: at most the plain MC error of 5 x 4e4 samples, ~2.2e-3
    # the value itself: the phase term 2 pi u alone moves it by more than 10 sigma
    wrong = math.cos(0.5 * sum(itg.params[1:])) * math.prod(
        math.sin(a / 2) / (a / 2) for a in itg.params[1:])
    if abs(wrong - itg.exact) > 20 * r["error"]:
        assert abs(r["result"] - wrong) > 10 * r["error"]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_run_closed_form_corner_peak(orc, prec, seed):
    """4-D corner peak (1 + sum a_i x_i)^-(d+1), a_i ~ U(0.5,3): the oracle
    integrator within 4 sigma of the inclusion-exclusion closed form, chi2/dof
    < 3, relative error below 1e-3.  The closed form itself is pinned by
    quadrature (test_closed_forms_of_library)."""
    from cuvegas_amd import integrands as I
    d, n, ns, N, iters = 4, 128, 6, 40000, 10
    rng = np.random.default_rng(40 + seed)
    itg = I.corner_peak(d, rng.uniform(0.5, 3.0, d).tolist())
    p, keep = orc.make_params(d, n, ns, N, iters, orc.F_CORNER, list(itg.params), seed=seed, precision=prec)
    r = orc.run(p, keep, ns ** d, snapshots=False)
    assert abs(r["result"] - itg.exact) < 4 * r["error"], (r["result"], itg.exact, r["error"])
    assert r["chi2_dof"] < 3.0
    assert r["error"] / itg.exact < 1e-3


def test_corner_peak_one_dimension_closed_form(orc):
    """d = 1: (1 + a x)^-2 integrates to 1 / (1 + a) over [0, 1]; the oracle's
    FP64 run at 2e5 samples per iteration gets it to 4 sigma and 2e-4."""
    a = 2.5
    p, keep = orc.make_params(1, 64, 50, 200000, 8, orc.F_CORNER, [a], seed=4)
    r = orc.run(p, keep, 50, snapshots=False)
    assert abs(r["result"] - 1.0 / (1.0 + a)) < 4 * r["error"]
    assert abs(r["result"] * (1.0 + a) - 1.0) < 2e-4


@pytest.mark.parametrize("prec", [0, 1])
def test_run_closed_form_gaussian_per_dimension_centres(orc, prec):
    """The batch family's Gaussian peak (a centre per dimension): 3-D, sigma
    0.12, centres near the faces, within 4 sigma of prod_i of the normal mass
    over [0, 1] and within 2e-3 relative.  One centre shared by every
    dimension changes the exact value by far more than that."""
    d, n, ns, N, iters = 3, 128, 8, 40000, 10
    sigma, c = 0.12, [0.15, 0.6, 0.8]
    norm = (1.0 / (sigma * math.sqrt(2 * math.pi))) ** d
    m = -1.0 / (2.0 * sigma * sigma)
    mass = [0.5 * (math.erf((1 - ci) / (sigma * math.sqrt(2))) + math.erf(ci / (sigma * math.sqrt(2))))
            for ci in c]
    exact = math.prod(mass)
    p, keep = orc.make_params(d, n, ns, N, iters, orc.F_GAUSSC, [norm, m, *c], seed=1, precision=prec)
    r = orc.run(p, keep, ns ** d, snapshots=False)
    assert abs(r["result"] - exact) < 4 * r["error"], (r["result"], exact, r["error"])
    assert abs(r["result"] / exact - 1) < 2e-3 and r["chi2_dof"] < 3.0
    shared = (0.5 * (math.erf(0.5 / (sigma * math.sqrt(2))) * 2)) ** d  # every centre at 0.5
    assert abs(shared - exact) > 20 * r["error"]


@pytest.mark.parametrize("prec", [0, 1])
def test_run_closed_form_mixture(orc, prec):
    """3-D four-component Gaussian mixture (weights, widths and centres drawn
    as for the autograd workload, seed 5): the oracle integrator, both forms,
    within 4 sigma of the erf closed form, chi2/dof < 3.  A width taken for a
    centre, a component weight dropped or a wrong normalisation moves the
    value by far more than 4 sigma."""
    from cuvegas_amd import integrands as I
    d, n, ns, N, iters = 3, 128, 6, 40000, 10
    th = I.mixture_theta(d, seed=5)
    exact, _ = I.mixture_exact(d, th)
    p, keep = orc.make_params(d, n, ns, N, iters, orc.F_MIXTURE, th, seed=1, precision=prec)
    r = orc.run(p, keep, ns ** d, snapshots=False)
    assert abs(r["result"] - exact) < 4 * r["error"], (r["result"], exact, r["error"])
    assert r["chi2_dof"] < 3.0
    assert r["error"] < 2e-3 * exact


def test_grad_replay_against_differences_and_closed_form(orc):
    """The oracle's replay of a frozen iteration of the 3-D mixture (uniform
    map and allocation of iteration 0): its value equals the forward's I_0
    (same samples); d value / d theta equals central differences of the
    replayed value itself (the samples do not move with theta, so the frozen
    estimator is smooth in theta) to 1e-5 of the largest component; and each
    component is within 4 of its own standard errors of the closed-form
    gradient.  A gradient term with a wrong sign, factor or index fails the
    differences."""
    from cuvegas_amd import integrands as I
    d, n, ns, N = 3, 64, 5, 50000
    th = np.array(I.mixture_theta(d, seed=5))
    p, keep = orc.make_params(d, n, ns, N, 1, orc.F_MIXTURE, th.tolist(), skip=0, seed=2)
    r = orc.run(p, keep, ns ** d, snapshots=False)
    C = ns ** d
    x0 = np.tile(np.linspace(0.0, 1.0, n + 1), (d, 1))
    nh0 = np.full(C, max(2, N // C), dtype=np.int32)
    I0, S20, g, gv = orc.grad_replay(p, keep, x0, nh0, 0, th)
    assert abs(I0 - r["hist_I"][0]) <= 1e-12 * abs(I0) and S20 > 0
    gmax = np.max(np.abs(g))
    for q in range(th.size):
        h = 1e-6 * max(abs(th[q]), 1e-3)
        tp, tm = th.copy(), th.copy()
        tp[q] += h
        tm[q] -= h
        fd = (orc.grad_replay(p, keep, x0, nh0, 0, tp)[0] - orc.grad_replay(p, keep, x0, nh0, 0, tm)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-5 * gmax, (q, fd, g[q])
    _, eg = I.mixture_exact(d, th.tolist())
    pull = (g - np.array(eg)) / np.sqrt(gv)
    assert np.all(np.abs(pull) < 4.0), pull

