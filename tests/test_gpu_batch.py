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
]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", BATCH_SHAPES, ids=["maxd6", "maxd16"])
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
    tasks, at two block sizes and task lengths: identical per-cube sums,
    allocation and first-iteration estimates, maps within 1e-12 (the bin
    atomics' order differs)."""
    from cuvegas_amd.batch import FAMILY_SOURCE, Batch, BatchProgram, serving_workload
    _say("batch-launch-shapes", prec)
    items, theta, _, _ = serving_workload(24, seed=9)
    items = [(d, min(N, 200_000), off) for d, N, off in items]
    ref = None
    for opts in ("block=256 task_windows=4", "block=256 task_windows=4 fill_grid=0",
                 "block=512 task_windows=2 fill_grid=0", "block=1024 task_windows=1"):
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
