"""Batched integrals: every item reproduces the single-integral path with
stream_id = item (bit-exact first iteration for a polynomial integrand), the
serving-style families agree with their closed forms, and graph replay agrees
with eager launches."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POLY = """
__device__ inline vg_real vg_poly(const vg_real* x, int d) {
    vg_real f = (vg_real)1;
    for (int i = 0; i < d; ++i) f = f * ((vg_real)6 * x[i] * ((vg_real)1 - x[i]));
    return f;
}
"""


def _say(*a):
    print("[gpu-case]", *a, flush=True)


BATCH_SHAPES = [
    # max_dim, n_intervals, items (d, n_eval_it, theta offset)
    (6, 64, [(2, 20_000, 0), (5, 60_000, 0), (3, 3_000, 0), (4, 250_000, 0), (1, 5_000, 0), (6, 40_000, 0)]),
    (16, 256, [(16, 50_000, 0), (9, 30_000, 0), (12, 20_000, 0)]),
    # a bin count that is not a power of two (map widths not pre-scaled)
    (5, 100, [(3, 30_000, 0), (5, 80_000, 0), (1, 4_000, 0), (4, 12_000, 0)]),
]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", BATCH_SHAPES, ids=["maxd6", "maxd16", "maxd5_nbins100"])
def test_batch_items_match_single_integrals(prec, shape):
    from cuvegas_amd import Integrator
    from cuvegas_amd.batch import Batch, BatchProgram
    max_dim, n, items = shape
    _say("batch-vs-single", prec, max_dim, n)
    bprog = BatchProgram(POLY + "__device__ vg_real integrand(vg_real* x, const vg_real* th, int d)"
                         " { return vg_poly(x, d); }", max_dim, n, prec)
    b = Batch(bprog, items, [0.0], max_it=4, skip=1, seed=7)
    b.run(1, use_graph=False)
    for i, (d, N, _) in enumerate(items):
        _, ns, C = b.item_info(i)
        single = Integrator(POLY + "__device__ vg_real integrand(vg_real* x) { return vg_poly(x, VG_NDIM); }",
                            d, N, n_intervals=n, n_strat=ns, max_it=4, skip=1, seed=7, stream_id=i,
                            precision=prec)
        single.run(1, use_graph=False)
        bj, bj2 = b.get_cube_sums(i)
        sj, sj2 = single.get_cube_sums()
        assert np.array_equal(bj, sj) and np.array_equal(bj2, sj2), i
        assert np.array_equal(b.get_nh(i), single.get_nh()), i
        assert np.max(np.abs(b.get_map(i) - single.get_map())) < 1e-12, i
        bI, bS2, bn = b.history(i)
        sI, sS2, sn = single.history()
        assert bI[0] == sI[0] and bS2[0] == sS2[0] and bn[0] == sn[0], i
        single.close()
    b.run(3)
    r = b.results()
    assert np.all(np.abs(r["result"] - 1.0) < 5 * r["error"])


def test_serving_families_against_closed_forms():
    from cuvegas_amd.batch import FAMILY_SOURCE, Batch, BatchProgram, serving_workload
    _say("batch-families")
    items, theta, exact, fams = serving_workload(48, seed=11)
    items = [(d, min(N, 300_000), off) for d, N, off in items]
    prog = BatchProgram(FAMILY_SOURCE, 12, 256, "fp32")
    b = Batch(prog, items, theta, max_it=8, skip=3, seed=1)
    b.run(8)
    r = b.results()
    pull = (r["result"] - exact) / r["error"]
    assert np.all(np.isfinite(r["result"])) and np.all(r["nevals"] > 0)
    # A Gaussian peak is resolvable only if uniform sampling puts some samples
    # in it: expected count N * prod(sigma sqrt(2 pi)) >= 20.  (A 12-D peak of
    # width 0.03 holds ~5e-14 of the cube: no sampler finds it with 1e5 points.)
    ok = np.ones(len(items), dtype=bool)
    for i, ((d, N, off), fam) in enumerate(zip(items, fams)):
        if fam == "gaussian":
            ok[i] = N * (theta[off + 1] * math.sqrt(2 * math.pi)) ** d >= 20
    assert ok.sum() >= 30
    assert np.mean(pull[ok] ** 2) < 3.0, pull[ok]
    assert np.max(np.abs(pull[ok])) < 5.0, pull[ok]


def test_batch_graph_matches_eager():
    from cuvegas_amd.batch import FAMILY_SOURCE, Batch, BatchProgram, serving_workload
    _say("batch-graph")
    items, theta, exact, _ = serving_workload(16, seed=5)
    items = [(d, min(N, 100_000), off) for d, N, off in items]
    prog = BatchProgram(FAMILY_SOURCE, 12, 256, "fp64")
    out = []
    for g in (True, False):
        b = Batch(prog, items, theta, max_it=4, skip=1, seed=2)
        b.run(4, use_graph=g)
        out.append((b.results(), [b.history(i)[0][0] for i in range(len(items))]))
        b.close()
    (ra, h0a), (rb, h0b) = out
    assert h0a == h0b  # first iteration: identical inputs, deterministic sums
    assert np.all(np.abs(ra["result"] - rb["result"]) <= 4 * np.hypot(ra["error"], rb["error"]))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_launch_shapes_agree(prec):
    """One block per task (the default) and a resident grid striding over the
    tasks, at three block sizes and task lengths, with 1, 2 and 4 copies of
    the LDS bins: identical per-cube sums, allocation and first-iteration
    estimates, maps within 1e-12 (the bin atomics' order differs)."""
    from cuvegas_amd.batch import FAMILY_SOURCE, Batch, BatchProgram, serving_workload
    _say("batch-launch-shapes", prec)
    items, theta, _, _ = serving_workload(24, seed=9)
    items = [(d, min(N, 200_000), off) for d, N, off in items]
    ref = None
    for opts in ("block=256 task_windows=4 bin_copies=1", "block=256 task_windows=4 fill_grid=0",
                 "block=512 task_windows=2 fill_grid=0", "block=512 task_windows=2 bin_copies=1",
                 "block=1024 task_windows=1", "block=1024 task_windows=1 bin_copies=4"):
        prog = BatchProgram(FAMILY_SOURCE, 12, 128, prec, options=opts)
        b = Batch(prog, items, theta, max_it=2, skip=0, seed=4)
        b.run(1, use_graph=False)
        got = [(b.get_cube_sums(i), b.get_nh(i), b.get_map(i), b.history(i)) for i in range(len(items))]
        b.close()
        if ref is None:
            ref = got
            continue
        for i, ((js, nh, x, h), (rjs, rnh, rx, rh)) in enumerate(zip(got, ref)):
            assert np.array_equal(js[0], rjs[0]) and np.array_equal(js[1], rjs[1]), (opts, i)
            assert np.array_equal(nh, rnh), (opts, i)
            assert np.max(np.abs(x - rx)) < 1e-12, (opts, i)
            assert h[0][0] == rh[0][0] and h[1][0] == rh[1][0] and h[2][0] == rh[2][0], (opts, i)


def test_batch_rejects_bad_items():
    """Items outside the program's limits or the parameter block are refused
    with an error before anything is launched."""
    from cuvegas_amd._lib import CuvegasError
    from cuvegas_amd.batch import Batch, BatchProgram
    _say("batch-bad-items")
    prog = BatchProgram(POLY + "__device__ vg_real integrand(vg_real* x, const vg_real* th, int d)"
                        " { return vg_poly(x, d); }", 4, 32, "fp64")
    for items, theta in [([(5, 1000, 0)], [0.0]),       # d > max_dim
                         ([(2, 1000, 1)], [0.0]),       # offset past the parameters
                         ([(2, 1000, 1)], []),          # offset with no parameters
                         ([(2, 0, 0)], [0.0])]:         # empty budget
        with pytest.raises(CuvegasError):
            Batch(prog, items, theta, max_it=2, skip=0)
    b = Batch(prog, [(2, 1000, 0)], [], max_it=2, skip=0)
    b.run(2)
    assert np.isfinite(b.results()["result"][0])
    b.close()


def _fixture_items():
    import json
    import os
    from cuvegas_amd.batch import workload_from_items
    with open(os.path.join(os.path.dirname(__file__), "golden", "workloads.json")) as f:
        return workload_from_items(json.load(f)["batch_512"])


def _resolvable(items, theta, fams):
    """A Gaussian peak is resolvable when uniform sampling expects at least 20
    points inside it: N * (sigma sqrt(2 pi))^d >= 20; the other families always are."""
    return np.array([fam != "gaussian" or N * (theta[off + 1] * math.sqrt(2 * math.pi)) ** d >= 20
                     for (d, N, off), fam in zip(items, fams)])


def _oracle_params(orc, items, theta, fams, i, n_strat, prec, K):
    """Oracle parameters of batch item i (Philox stream i, seed 0)."""
    d, N, off = items[i]
    fam = fams[i]
    if fam == "gaussian":
        fid, par = orc.F_GAUSSC, theta[off + 2:off + 4 + d]
    elif fam == "oscillatory":
        fid, par = orc.F_OSC, theta[off + 1:off + 2 + d]
    else:
        fid, par = orc.F_CORNER, theta[off + 1:off + 1 + d]
    return orc.make_params(d, 256, n_strat, N, 1, fid, list(par), skip=0, K=K, seed=0, stream=i,
                           precision=0 if prec == "fp64" else 1, threads=8)


def _oracle_item(orc, items, theta, fams, i, n_strat, prec, K):
    """Oracle first iteration of batch item i."""
    p, keep = _oracle_params(orc, items, theta, fams, i, n_strat, prec, K)
    return orc.run(p, keep, n_strat ** items[i][0])


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_batch512_full_configuration(orc, prec):
    """The 512-integral serving workload at its configuration, read from the
    committed fixture (d 2..12, Gaussian / oscillatory / corner peak, budgets
    1e4..1e7 per iteration, 256 bins, 10 iterations in one hipGraph replay, 5
    averaged).  After the first iteration, one item of each family (the
    smallest budget of at least 1e5 with d >= 4, resolvable) is checked against the
    oracle's first iteration on the same Philox stream: bin-exact allocation
    and per-cube sums for the Gaussian (FP64: the project's exp on both sides)
    and the corner peak; the oscillatory per-cube sums within the bound that
    cos' ulp differences allow.  At the end: invariants of every item and the
    pulls of the resolvable items against their closed forms."""
    from cuvegas_amd.batch import FAMILY_SOURCE, Batch, BatchProgram
    items, theta, exact, fams = _fixture_items()
    _say("batch512-full", prec)
    assert len(items) == 512
    K = 16
    prog = BatchProgram(FAMILY_SOURCE, 12, 256, prec, run_len=K)
    b = Batch(prog, items, theta, max_it=10, skip=5, seed=0)
    b.run(1, use_graph=True)
    ok = _resolvable(items, theta, fams)
    for fam in ("gaussian", "oscillatory", "corner"):
        cand = [i for i, ((d, N, _), f) in enumerate(zip(items, fams))
                if f == fam and d >= 4 and N >= 100_000 and ok[i]]
        i = min(cand, key=lambda k: items[k][1])
        _say("batch512-item", prec, fam, i, items[i])
        d, N, _ = items[i]
        _, ns, C = b.item_info(i)
        o = _oracle_item(orc, items, theta, fams, i, ns, prec, K)
        JF, JF2 = b.get_cube_sums(i)
        I0, S20, ne0 = b.history(i)
        assert ne0[0] == o["hist_n"][0]
        exact_sums = fam == "corner" or (fam == "gaussian" and prec == "fp64")
        if exact_sums:
            assert np.array_equal(JF, o["JF_first"]) and np.array_equal(JF2, o["JF2_first"]), (fam, i)
            assert np.array_equal(b.get_nh(i), o["nh_first"]), (fam, i)
            assert I0[0] == o["hist_I"][0] and S20[0] == o["hist_S2"][0]
        else:
            # per sample, J f differs by a few ulps of the FP32 value (relative
            # 4e-7) or, for the FP64 cos, by 2 ulps of 1 (|f| <= 1, J = 1 under
            # the initial uniform map); the FP64 sums of n values differ by at
            # most 2 (n - 1) u sum|x| from their rounding plus those differences
            nh0 = np.full(C, max(2, int(N / C)), dtype=np.int32)
            p, keep = _oracle_params(orc, items, theta, fams, i, ns, prec, K)
            x0 = np.tile(np.linspace(0.0, 1.0, 257), (d, 1))
            _, _, _, ojf = orc.sample_dump(p, keep, x0, nh0, 0, int(nh0.sum()))
            start = np.concatenate([[0], np.cumsum(nh0)[:-1]])
            A = np.add.reduceat(np.abs(ojf), start)
            dv = 4e-7 * A if prec == "fp32" else 4.5e-16 * nh0
            bound1 = 2.3e-16 * nh0 * A + dv
            assert np.all(np.abs(JF - o["JF_first"]) <= bound1), (fam, i)
            assert abs(I0[0] - o["hist_I"][0]) <= 2 * math.pow(1.0 / ns, d) * np.sum(bound1 / nh0) + \
                1e-15 * abs(o["hist_I"][0])
    b.run(9, use_graph=True)
    r = b.results()
    b_nh = [b.get_nh(i) for i in range(512)]
    infos = [b.item_info(i) for i in range(512)]
    for i, ((d, N, _), (_, ns, C)) in enumerate(zip(items, infos)):
        assert b_nh[i].min() >= 2 and b_nh[i].sum() <= N + 2 * C, i
        I_k, S2, ne = b.history(i)
        assert len(ne) == 10 and r["nevals"][i] == ne.sum(), i
    for i in (0, 255, 511):
        x = b.get_map(i)
        assert np.all(np.diff(x, axis=1) > 0) and np.all(x[:, 0] == 0) and np.all(x[:, -1] == 1)
    b.close()
    assert np.all(np.isfinite(r["result"])) and np.all(r["error"] > 0)
    pull = (r["result"] - exact) / r["error"]
    assert ok.sum() >= 400
    assert np.sqrt(np.mean(pull[ok] ** 2)) < 1.6, np.sqrt(np.mean(pull[ok] ** 2))
    assert np.mean(np.abs(pull[ok]) < 3) > 0.95


def test_batch_timing_of_eager_runs():
    """cvg_batch_set_timing / get_timing: an eager run records one (fill,
    iteration) pair per iteration with 0 < fill <= iteration; graph replays
    record none, and the results are those of an untimed run."""
    from cuvegas_amd.batch import FAMILY_SOURCE, Batch, BatchProgram
    items, theta, exact, fams = _fixture_items()
    items = items[:24]
    prog = BatchProgram(FAMILY_SOURCE, 12, 256, "fp32", run_len=16)
    _say("batch-timing")
    a = Batch(prog, items, theta, max_it=6, skip=2, seed=0)
    a.set_timing(True)
    a.run(3, use_graph=False)
    f, t = a.timing()
    assert len(f) == 3 and np.all(f > 0) and np.all(f <= t)
    a.run(3, use_graph=True)
    f2, _ = a.timing()
    assert len(f2) == 3
    ra = a.results()
    a.close()
    b = Batch(prog, items, theta, max_it=6, skip=2, seed=0)
    b.run(3, use_graph=False)
    b.run(3, use_graph=True)
    rb = b.results()
    b.close()
    assert np.array_equal(ra["nevals"], rb["nevals"])
    assert np.allclose(ra["result"], rb["result"], rtol=1e-9, atol=0)

