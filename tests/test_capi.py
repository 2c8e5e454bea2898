"""C ABI: the library loads, exports every declared symbol, validates its
arguments and compiles integrands with hipRTC -- all without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "cuvegas.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cvg_[a-z_0-9]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    from cuvegas_amd import _lib
    names = _declared()
    assert len(names) >= 30
    raw = C.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(raw, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def _params(**kw):
    from cuvegas_amd.api import make_params
    base = dict(n_dim=3, n_intervals=64, n_strat=4, max_it=5, skip=1, n_eval_it=1000,
                alpha=0.5, beta=0.75, seed=1)
    base.update(kw)
    return make_params(**base)[0]


@pytest.mark.parametrize("bad", [dict(n_dim=0), dict(n_dim=65), dict(n_intervals=1),
                                 dict(n_strat=0), dict(n_strat=2**16, n_dim=2),
                                 dict(precision=7), dict(n_strat=70000, n_intervals=70000, n_dim=1)])
def test_invalid_shapes_are_rejected(bad):
    from cuvegas_amd._lib import CVG_ERR_INVALID, lib
    h = C.c_void_p()
    rc = lib.cvg_program_create(b"__device__ vg_real integrand(vg_real* x){return x[0];}",
                                C.byref(_params(**bad)), None, C.byref(h))
    assert rc == CVG_ERR_INVALID
    assert lib.cvg_last_error()


def test_compile_error_reports_log():
    from cuvegas_amd import CuvegasError
    from cuvegas_amd.api import Program
    with pytest.raises(CuvegasError) as ei:
        Program("__device__ vg_real integrand(vg_real* x) { return undefined_symbol(x); }", _params())
    assert ei.value.code == -2
    assert "undefined_symbol" in str(ei.value)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_library_integrands_compile(prec):
    from cuvegas_amd import integrands as I
    from cuvegas_amd.api import Program
    for name in sorted(I.LIBRARY):
        prog = Program(I.by_name(name, 3).source, _params(precision=prec))
        assert prog.code_size > 1000, name
        assert prog.code()[:4] == b"\x7fELF"


def test_integrand_forms_compile():
    """config.cuh form, extension status form and the parametrised form."""
    from cuvegas_amd.api import Program
    forms = [
        "__host__ __device__ double integrand(double* x) { return x[0] * x[1]; }\nconst int n_dim = 3;",
        'extern "C" __device__ int integrand(double* ret, double* x) { *ret = x[0]; return 0; }',
        "__device__ vg_real integrand(vg_real* x, const vg_real* t) { return t[0] * x[0]; }",
    ]
    for src in forms:
        assert Program(src, _params()).code_size > 0


def test_state_needs_device():
    from cuvegas_amd._lib import CVG_ERR_NODEVICE, lib
    from cuvegas_amd.api import Program
    ndev = C.c_int(0)
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    p = _params()
    prog = Program("__device__ vg_real integrand(vg_real* x){return x[0];}", p)
    h = C.c_void_p()
    assert lib.cvg_state_create(prog.handle, C.byref(p), C.byref(h)) == CVG_ERR_NODEVICE
    del ndev


def test_reference_entry_points_importable():
    from vegas_cuda_extension.cuda_extension import integrate_cuda  # noqa: F401
    from vegas_cuda_extension.python_wrapper import integrate
    import inspect
    sig = inspect.signature(integrate)
    names = list(sig.parameters)[:17]
    assert names == ["device_function", "n_dim", "domain", "n_evals", "max_it", "skip", "batch_size",
                     "n_intervals", "n_strat", "alpha", "beta", "seed", "multi_gpu", "multi_cpu",
                     "gpu_num", "cpu_num", "gpu_thr"]
    assert sig.parameters["n_evals"].default == 1000000
    assert sig.parameters["max_it"].default == 20
    assert sig.parameters["n_intervals"].default == 1024


def test_default_n_strat_matches_reference_formula():
    from cuvegas_amd.api import default_n_strat
    assert default_n_strat(10**7, 10, 4) == 22
    assert default_n_strat(2 * 10**9, 20, 8) == 8
    assert default_n_strat(5 * 10**7, 10, 6) == 10


def test_closed_forms_of_library():
    import math
    from cuvegas_amd import integrands as I
    assert I.ridge(4).exact == pytest.approx(0.851317758241297887, rel=1e-12)
    assert I.gaussian_peak(4).exact == pytest.approx(0.99999770678939702466, rel=1e-13)
    assert I.double_gaussian(8).exact == pytest.approx(0.99999999983553688726, rel=1e-13)
    assert I.gaussian(4).exact == pytest.approx(1.0, rel=1e-15)
    # corner peak: inclusion-exclusion vs. 1-D quadrature in d = 1
    assert I.corner_peak(1, [2.0]).exact == pytest.approx(1.0 / 3.0, rel=1e-14)  # int (1+2x)^-2
    from scipy import integrate as si
    q, _ = si.nquad(lambda x, y: (1 + 0.5 * x + 1.5 * y) ** -3, [[0, 1], [0, 1]])
    assert I.corner_peak(2, [0.5, 1.5]).exact == pytest.approx(q, rel=1e-10)
    assert I.oscillatory(1, 0.0, [1.0]).exact == pytest.approx(math.sin(1.0), rel=1e-14)
    # the reference library's closed forms (src/commons/functions.cu) against
    # 2-D quadrature of their definitions
    quad2 = {
        "linear": lambda x, y: x + y,
        "roosarnold": lambda x, y: (abs(4 * x - 2) + abs(4 * y - 2)) / 2,
        "morokoffcalfisch": lambda x, y: (x * y) ** 0.5 * 1.5 ** 2,
        "expon": lambda x, y: math.exp(-(x * x + y * y)),
        "cosin": lambda x, y: math.cos(x) * math.cos(y),
        "dirichlet": lambda x, y: x ** 4 * y ** 4 / (math.gamma(5) ** 2 / math.gamma(10)),
        "sinexp": lambda x, y: math.sin(x) + math.exp(y),
        "one": lambda x, y: 1.0,
    }
    for name, f in quad2.items():
        q, _ = si.nquad(f, [[0, 1], [0, 1]], opts={"points": [0.5]})
        assert I.by_name(name, 2).exact == pytest.approx(q, rel=1e-9), name
    assert I.hellekalek(2).exact == 0.0


def test_run_len_must_be_power_of_two():
    from cuvegas_amd._lib import CVG_ERR_INVALID, lib
    for bad in (3, 6, 5000):
        h = C.c_void_p()
        rc = lib.cvg_program_create(b"__device__ vg_real integrand(vg_real* x){return x[0];}",
                                    C.byref(_params(run_len=bad)), None, C.byref(h))
        assert rc == CVG_ERR_INVALID


@pytest.mark.parametrize("opts,max_dim,nbins", [("bin_copies=3", 12, 256), ("bin_copies=0", 4, 64),
                                                ("bin_copies=4", 12, 1024), ("block=96", 4, 64)])
def test_batch_launch_options_are_checked(opts, max_dim, nbins):
    """The batched program refuses a bin-copy count other than 1, 2 or 4, copies
    whose bins do not fit in LDS (4 x 12 dimensions x 1024 bins x 12 bytes), and
    a block that is not a multiple of 64, before it compiles anything."""
    from cuvegas_amd._lib import CVG_ERR_INVALID, lib
    h = C.c_void_p()
    rc = lib.cvg_batch_program_create(b"__device__ vg_real integrand(vg_real* x){return x[0];}", max_dim,
                                      C.byref(_params(n_dim=max_dim, n_intervals=nbins)), opts.encode(),
                                      C.byref(h))
    assert rc == CVG_ERR_INVALID
    assert not h.value


def test_program_cache_reuses_compiled_programs():
    """integrate() reuses the hipRTC program of an identical (source, shape,
    options) and compiles a new one when any of them changes; the cache is
    bounded (least recently used dropped)."""
    from cuvegas_amd import api
    src = "__device__ vg_real integrand(vg_real* x) { return x[0] + x[1]; }"
    api._PROGRAMS.clear()
    p1, _ = api.make_params(3, 64, 4, 5, 1, 10_000, 0.5, 0.75, 0)
    a = api.cached_program(src, p1)
    assert api.cached_program(src, p1) is a
    p2, _ = api.make_params(3, 64, 5, 5, 1, 10_000, 0.5, 0.75, 0)  # other n_strat
    assert api.cached_program(src, p2) is not a
    assert api.cached_program(src, p1, "pipe=0") is not a
    p3, _ = api.make_params(3, 64, 4, 9, 2, 50_000, 0.3, 0.5, 7)  # loop parameters only: same program
    assert api.cached_program(src, p3) is a
    for k in range(api._PROGRAMS_MAX + 2):
        pk, _ = api.make_params(2, 16 + k, 3, 2, 1, 1_000, 0.5, 0.75, 0)
        api.cached_program(src, pk)
    assert len(api._PROGRAMS) == api._PROGRAMS_MAX
    assert all(v is not a for v in api._PROGRAMS.values())
    api._PROGRAMS.clear()


def test_batch_workload_fixture_matches_its_draw():
    """The committed 512-integral workload is the seed-3 draw of the serving
    family generator, and its items and parameter blocks are built as the
    batched integrand reads them."""
    import json
    import os
    from cuvegas_amd.batch import draw_serving_items, workload_from_items
    with open(os.path.join(os.path.dirname(__file__), "golden", "workloads.json")) as f:
        fixture = json.load(f)["batch_512"]
    assert draw_serving_items(512, 3) == fixture
    items, theta, exact, fams = workload_from_items(fixture)
    assert len(items) == 512 and set(fams) == {"gaussian", "oscillatory", "corner"}
    for (d, N, off), fam, e in zip(items, fams, fixture):
        assert d == e["d"] and N == e["neval"] and 2 <= d <= 12 and 10_000 <= N <= 10_000_000
        assert theta[off] == {"gaussian": 0.0, "oscillatory": 1.0, "corner": 2.0}[fam]
    assert np.all(np.isfinite(exact))


def test_roofline_slot_counts():
    """The algorithmic work per evaluation of the benchmark configurations
    (149, 304, 528 and 302 VALU lane-slots for the 4-D Gaussian, the 8-D
    double Gaussian, the 16-D oscillatory and the 6-D mixture) and per
    replayed sample of the mixture's 32-parameter backward (302 - 12 + 68 +
    96 = 454), and the MI355X issue peak (256 CUs x 64 lanes x 2.4 GHz)."""
    from cuvegas_amd.roofline import slots_per_eval, slots_per_replayed_sample, valu_peak
    assert slots_per_eval(4, "gaussian_peak") == 149
    assert slots_per_eval(8, "double_gaussian") == 304
    assert slots_per_eval(16, "oscillatory") == 528
    assert slots_per_eval(6, "gaussian_mixture") == 302
    assert slots_per_replayed_sample(6, "gaussian_mixture", 32) == 454
    assert valu_peak(256, 2.4e9) == pytest.approx(3.93216e13)


def test_lds_roofline_and_rate_arguments():
    """The LDS roofline of a sampling launch: d bin updates (a ds_add_f64 and
    a ds_add_u32) per evaluation against the CUs' rate for the pair; 1.33e8
    evaluations of an 8-D integrand in 1.25 ms at 16.7 cycles per instruction
    on 256 CUs at 2.4 GHz is 8.52e11 of 1.177e12 updates/s.  cvg_device_rates
    refuses an output array shorter than CVG_RATE_COUNT before touching a
    device."""
    from cuvegas_amd._lib import CVG_ERR_INVALID, lib
    from cuvegas_amd.roofline import RATE_NAMES, lds_roofline
    r = lds_roofline(1.33e8, 8, 1.25e-3, 256, 2.4e9, 16.7)
    assert r["achieved"] == pytest.approx(1.33e8 * 8 / 1.25e-3 / 1e9)
    assert r["peak"] == pytest.approx(256 * 2.4e9 * 64 / 33.4 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / r["peak"])
    assert len(RATE_NAMES) == 9
    buf = (C.c_double * 8)()
    assert lib.cvg_device_rates(buf, 8) == CVG_ERR_INVALID
    assert b"CVG_RATE_COUNT" in lib.cvg_last_error()


def _run_main(*args, cwd=None):
    import subprocess
    exe = os.path.join(ROOT, "cuvegas_amd", "cuvegas_main")
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=120, cwd=cwd)


def test_standalone_main_reads_reference_config():
    """cuvegas_main (the reference's standalone src/main.cu caller) reads a
    configuration header in the reference's format: the settings from their
    initialisers, the reference's sizing (n_eval_it = tot_eval / max_it,
    n_strat = (int)(n_eval_it / 4)^(1/d): 1e7 / 20 in 4-D gives 18 strata,
    104,976 cubes) and the integrand, which hipRTC compiles with the library
    standing in for commons/functions.cuh (no GPU needed)."""
    r = _run_main(os.path.join(ROOT, "tests", "golden", "config_ridge4.cuh"), "--compile-only")
    assert r.returncode == 0, r.stderr
    lines = dict(l.split(": ", 1) for l in r.stdout.splitlines() if ": " in l)
    assert lines["Strat"] == "18" and lines["Cubes"] == "104976" and lines["Intervals"] == "1024"
    assert lines["Evaluations per iteration"] == "500000" and lines["Iterations"] == "20 (skip 5)"
    assert lines["Alpha"] == "0.5" and lines["Beta"] == "0.75" and lines["Seed"] == "123"


def test_standalone_main_refuses_bad_configs(tmp_path):
    """An include other than the integrand library or the C maths/I-O headers,
    a missing n_dim, unusable settings, a setting initialised by an
    expression (not evaluated here) and an integrand hipRTC rejects are
    reported on stderr with a non-zero status."""
    cases = {
        "inc.cuh": '#include "mystuff.h"\nconst int n_dim = 2;\n__host__ __device__ double integrand(double* x) { return 1.0; }\n',
        "nodim.cuh": "__host__ __device__ double integrand(double* x) { return 1.0; }\n",
        "skip.cuh": "int max_it = 5;\nint skip = 5;\nconst int n_dim = 2;\n__host__ __device__ double integrand(double* x) { return 1.0; }\n",
        "bad.cuh": "const int n_dim = 2;\n__host__ __device__ double integrand(double* x) { return nope(x); }\n",
        "expr.cuh": "long int tot_eval = 10 * 1000000;\nconst int n_dim = 2;\n__host__ __device__ double integrand(double* x) { return 1.0; }\n",
    }
    for name, text in cases.items():
        p = tmp_path / name
        p.write_text(text)
        r = _run_main(str(p), "--compile-only")
        assert r.returncode != 0 and r.stderr.strip(), (name, r.stdout)
    assert "unsupported include" in _run_main(str(tmp_path / "inc.cuh"), "--compile-only").stderr
    assert "nope" in _run_main(str(tmp_path / "bad.cuh"), "--compile-only").stderr
    assert "tot_eval is not a plain number" in _run_main(str(tmp_path / "expr.cuh"), "--compile-only").stderr


def test_standalone_integrand_library_closed_forms(tmp_path):
    """vg_functions.h (the reference's integrand library as the standalone
    program compiles it, functions.cu:5-111) built as host code and
    integrated by the midpoint rule against each function's closed form in
    1-D (every function but sinexp) and 2-D.  A wrong constant, sign or
    operand shows up as a missed integral."""
    import math
    import subprocess
    src = tmp_path / "fn.cpp"
    src.write_text("""
#define __host__
#define __device__
#include <math.h>
#include "vg_functions.h"
typedef double (*fn)(double*, int);
static const fn table[] = {linear, roosarnold, morokoffcalfisch, hellekalek, ridge, expon, cosin,
                           dirichlet, gaussian, sinexp, one};
extern "C" double vg_quad(int k, int d, int m) {  // midpoint rule, m points per dimension
    double x[8], s = 0.0;
    long long n = 1;
    for (int i = 0; i < d; ++i) n *= m;
    for (long long c = 0; c < n; ++c) {
        long long t = c;
        for (int i = 0; i < d; ++i) { x[i] = ((double)(t % m) + 0.5) / m; t /= m; }
        s += table[k](x, d);
    }
    return s / (double)n;
}
""")
    so = tmp_path / "libfn.so"
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-I", os.path.join(ROOT, "cuvegas_amd", "csrc"),
                    str(src), "-o", str(so)], check=True)
    q = C.CDLL(str(so)).vg_quad
    q.restype, q.argtypes = C.c_double, [C.c_int, C.c_int, C.c_int]
    names = ["linear", "roosarnold", "morokoffcalfisch", "hellekalek", "ridge", "expon", "cosin",
             "dirichlet", "gaussian", "sinexp", "one"]
    erfm = lambda mu, sg: 0.5 * (math.erf((1 - mu) / (sg * math.sqrt(2))) + math.erf(mu / (sg * math.sqrt(2))))

    def ridge_exact(d):
        return (100 / math.pi) ** 2 / 1000 * sum(
            (math.sqrt(math.pi / 100) * 0.5 * (math.erf(10 * (1 - k / 999)) + math.erf(10 * k / 999))) ** d
            for k in range(1000))

    exact = {
        "linear": lambda d: d / 2, "roosarnold": lambda d: 1.0, "morokoffcalfisch": lambda d: 1.0,
        "hellekalek": lambda d: 0.0, "ridge": ridge_exact,
        "expon": lambda d: (math.sqrt(math.pi) / 2 * math.erf(1)) ** d, "cosin": lambda d: math.sin(1) ** d,
        "dirichlet": lambda d: 0.2 ** d * math.gamma(5 * d) / math.gamma(5) ** d,
        "gaussian": lambda d: erfm(0.5, 0.01) ** d, "sinexp": lambda d: (1 - math.cos(1)) + (math.e - 1),
        "one": lambda d: 1.0}
    for k, name in enumerate(names):
        for d, m in ((1, 20000 if name == "ridge" else 200000), (2, 300 if name == "ridge" else 1000)):
            if name == "sinexp" and d == 1:
                continue
            got, want = q(k, d, m), exact[name](d)
            tol = 2e-3 if name in ("morokoffcalfisch", "gaussian") and d == 2 else 1e-4
            assert abs(got - want) <= tol * max(1.0, abs(want)), (name, d, got, want)
    # the closed form used above gives the reference's default run its value
    # (4-D ridge, 0.851317758241297887)
    assert ridge_exact(4) == pytest.approx(0.851317758241297887, rel=1e-13)


def test_combine_rule():
    """cuvegas_amd.combine: the reference's inverse-variance mean of the
    iterations >= skip (cuda_wrapper.cu:758-766) with chi2/dof, checked on
    hand-computed values; iterations of zero variance are exact and give
    their mean with error 0."""
    import math
    from cuvegas_amd import combine
    I = [9.0, 1.0, 2.0, 4.0]
    S2 = [1.0, 1.0, 4.0, 4.0]
    res, err, chi2 = combine(I, S2, 1)
    # weights 1, 1/4, 1/4: mean (1 + 0.5 + 1) / 1.5, error 1 / sqrt(1.5)
    assert res == pytest.approx(2.5 / 1.5, rel=1e-15)
    assert err == pytest.approx(1.0 / math.sqrt(1.5), rel=1e-15)
    m = 2.5 / 1.5
    assert chi2 == pytest.approx(((1 - m) ** 2 + (2 - m) ** 2 / 4 + (4 - m) ** 2 / 4) / 2, rel=1e-14)
    assert combine([5.0, 3.0, 3.5, 7.0], [1.0, 0.0, 0.0, 2.0], 1) == (3.25, 0.0, 0.0)

