"""The reference's native entry point, integrate_cuda, over the C ABI.

Same 17 positional arguments and 14-tuple result as the pybind binding
(vegas_cuda/src/vegas_cuda_extension/cuda_wrapper.cu:290, :869, :874-876).
The 16th argument carries the integrand's HIP C++ source instead of PTX; the
17th (the kernel PTX path in the reference) carries options ("fp32" selects
the FP32 form; other tokens are passed to hipRTC).
"""
from __future__ import annotations

import ctypes as C

from ._lib import check, cvg_ref_result, lib


def integrate_cuda(tot_eval, max_it, skip, max_batch_size, n_intervals, n_strat, n_dim, alpha,
                   beta, seed, multi_gpu, multi_cpu, gpu_num, cpu_num, gpu_thr, integrand_src,
                   options=""):
    out = cvg_ref_result()
    check(lib.cvg_integrate_cuda(int(tot_eval), int(max_it), int(skip), int(max_batch_size),
                                 int(n_intervals), int(n_strat), int(n_dim), float(alpha),
                                 float(beta), int(seed) & 0xFFFFFFFF, int(bool(multi_gpu)),
                                 int(bool(multi_cpu)), int(gpu_num), int(cpu_num), float(gpu_thr),
                                 integrand_src.encode(), (options or "").encode(), C.byref(out)))
    return (out.res, out.err, out.n_strat, out.n_cubes, out.n_intervals, out.tot_nevals,
            out.tot_time, out.avg_it_time, out.seed, bool(out.multi_gpu), bool(out.multi_cpu),
            out.gpu_num, out.omp_threads_num, out.gpu_thr)
