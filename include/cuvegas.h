/*
 * cuvegas.h -- C ABI of the MI355X VEGAS+ integrator (libcuvegas_amd.so).
 *
 * Plain pointers and sizes only.  Every function returns CVG_OK (0) or a
 * negative status; cvg_last_error() gives the message of the calling thread's
 * last failure.  Nothing here calls exit() (the reference exits the process on
 * any CUDA error: vegas_cuda/cuda-samples/Common/helper_cuda.h:582-595,
 * vegas_cuda/src/vegas_cuda_extension/commons/common.cuh:39-45).
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   cvg_integrate_cuda   <- integrate(...) bound as integrate_cuda
 *                           (vegas_cuda/src/vegas_cuda_extension/cuda_wrapper.cu:290,
 *                            pybind at :874-876; 14-tuple at :869)
 *   cvg_program_create   <- jit(): cuLink of cuda_kernel.ptx + integrand PTX
 *                           (cuda_wrapper.cu:221-286; python_wrapper.py:49-59)
 *   cvg_state_create     <- allocation and initialisation (cuda_wrapper.cu:348-543)
 *   cvg_state_run        <- the iteration loop (cuda_wrapper.cu:575-773)
 *   cvg_state_result     <- weighted average and error (cuda_wrapper.cu:750-766,869)
 *   cvg_state_refine     <- map refinement (cuda_wrapper.cu:719-746)
 *   cvg_state_reallocate <- stratification reallocation (cuda_wrapper.cu:729-730)
 *   cvg_state_dump       <- no counterpart: diagnostics of the sampling step
 *   cvg_state_get_..., cvg_state_set_... <- no counterpart: the reference keeps all state per call
 *                           (cuda_wrapper.cu:394-450,779-851)
 *
 * Threading: handles are independent; a state must not be used from two
 * threads at once.  All device work of a state is ordered on its stream.
 */
#ifndef CUVEGAS_H
#define CUVEGAS_H

#ifdef __cplusplus
extern "C" {
#endif

#define CVG_OK 0
#define CVG_ERR_INVALID (-1)   /* bad argument or unsupported shape */
#define CVG_ERR_COMPILE (-2)   /* hipRTC rejected the integrand; see cvg_program_log */
#define CVG_ERR_HIP (-3)       /* HIP runtime error (message names the call) */
#define CVG_ERR_ALLOC (-4)     /* device or host allocation failed */
#define CVG_ERR_CAPACITY (-5)  /* allocation or iteration count exceeded the buffers */
#define CVG_ERR_NODEVICE (-6)  /* no HIP device visible */

#define CVG_FP64 0   /* FP64 sampling, map lookup and integrand */
#define CVG_FP32C 1  /* FP32 sampling, map lookup and integrand; FP64 sums and map */

typedef struct cvg_program cvg_program;
typedef struct cvg_state cvg_state;

typedef struct cvg_params {
    int n_dim;            /* dimensions d (1..64) */
    int n_intervals;      /* map bins per dimension (>= 2) */
    int n_strat;          /* strata per dimension (>= 1), n_strat^d < 2^31 */
    int max_it;           /* iterations to run (history capacity) */
    int skip;             /* leading iterations left out of the weighted average */
    int precision;        /* CVG_FP64 or CVG_FP32C */
    long long n_eval_it;  /* sample budget per iteration */
    double alpha;         /* map damping exponent */
    double beta;          /* stratification exponent: n_h ~ (sigma_h^2)^beta */
    unsigned int seed;    /* Philox key word 0 */
    unsigned int stream_id; /* Philox key word 1 (separates independent integrals) */
    int run_len;          /* samples per lane run; 0 selects 16 */
    int has_domain;       /* non-zero: domain[] below is used */
    const double* domain; /* 2*n_dim values: lower bounds, then ranges */
    int n_theta;          /* parameters of a parametrised integrand (0 if none) */
} cvg_params;

/* Result of a run; timings are wall clock (ms). */
typedef struct cvg_result {
    double result;        /* inverse-variance weighted mean over iterations >= skip */
    double error;         /* 1 / sqrt(sum 1/sigma_k^2) */
    double chi2_dof;      /* sum (I_k - result)^2 / sigma_k^2 / (K - 1) */
    long long tot_nevals; /* integrand evaluations over all iterations */
    double loop_ms;       /* device time of the iteration loop (HIP events) */
    double compile_ms;    /* hipRTC compile and module load */
    double setup_ms;      /* allocation and initialisation */
    int iterations;       /* iterations completed */
    int n_cubes;
    int n_strat;
    int n_intervals;
    /* No reference counterpart: the same weighted mean's error and chi2/dof
     * with the unbiased per-cube variance sigma_h^2 / (n_h - 1) in place of the
     * reference's sigma_h^2 / n_h (which is low by n_h / (n_h - 1), a factor
     * of 2 for the many cubes holding the minimum of 2 samples). */
    double error_unbiased;
    double chi2_dof_unbiased;
} cvg_result;

/* The 14 values the reference's integrate_cuda returns (cuda_wrapper.cu:869). */
typedef struct cvg_ref_result {
    double res;
    double err;
    int n_strat;
    int n_cubes;
    int n_intervals;
    long long tot_nevals;
    double tot_time;      /* ms, whole call */
    double avg_it_time;   /* ms, loop time / max_it */
    unsigned int seed;
    int multi_gpu;
    int multi_cpu;
    int gpu_num;
    int omp_threads_num;
    double gpu_thr;
} cvg_ref_result;

/* Last error message of the calling thread ("" if none). */
const char* cvg_last_error(void);

/* Device facts used for roofline figures. */
int cvg_device_info(int* n_cu, int* clock_khz, char* arch, int arch_len);

/* Measured issue costs of the sampling kernel's instructions on the current
 * device, in cycles of the nominal clock per wave64 instruction (CU cycles for
 * the LDS, SIMD cycles for the VALU): the ceilings its roofline is read
 * against.  No counterpart in the reference (it reports no roofline).  About
 * 30 ms of probe kernels on a private stream; allocates 4 MB for their output.
 * n >= CVG_RATE_COUNT; indices: */
#define CVG_RATE_LDS_F64_RANDOM 0 /* ds_add_f64, 64 lanes at random slots of 1024 */
#define CVG_RATE_LDS_U32_RANDOM 1 /* ds_add_u32, the same slots */
#define CVG_RATE_LDS_PAIR_RANDOM 2 /* one of each per slot (the bin update), per instruction */
#define CVG_RATE_LDS_F64_SEQ 3    /* ds_add_f64, 64 consecutive slots: no bank conflicts */
#define CVG_RATE_LDS_U32_SEQ 4    /* ds_add_u32, 64 consecutive slots */
#define CVG_RATE_VALU_F32_FMA 5   /* v_fma_f32 */
#define CVG_RATE_VALU_F64_FMA 6   /* v_fma_f64 */
#define CVG_RATE_VALU_U32_ADD 7   /* v_add_u32 and v_lshrrev_b32, per instruction */
#define CVG_RATE_VALU_MAD_U64 8   /* v_mad_u64_u32 (Philox's 32x32 -> 64 product) */
#define CVG_RATE_COUNT 9
int cvg_device_rates(double* cycles, int n);

/* Compile an integrand source (HIP C++) for the given shape with hipRTC.
 * shape supplies n_dim, n_intervals, n_strat, precision, run_len, has_domain.
 * options: space-separated launch knobs and extra hipRTC options, or NULL.
 * Knobs (consumed here): block=64..1024 (multiple of 64), pipe=0|1 (one
 * sample per lane step; 1: the next sample's map entries are requested before
 * the current one is evaluated, the default for n_dim <= 8), pair=0|1 (two
 * samples per step), lds_bins=0|1 (1, the default: per-block bins in LDS,
 * in one sampling pass per group of dimensions when d x n_intervals bins do
 * not fit; 0: every sample to the global bins), lds_map=0|1 (FP32 map table
 * in LDS), bin32=0|1, bin_copies=1..8 and copy_dim0=i (copies of the LDS bins
 * of dimensions i.. that lanes add into by lane mod copies; by default 8, 4
 * or 2 copies, as LDS allows, of the dimensions whose strata a window's cubes
 * share when a stratum spans at most 32 bins), grad=0|1 (compile the backward
 * replay kernel).  Every other token is passed to hipRTC.
 * On CVG_ERR_COMPILE, *out is still set so cvg_program_log can be read. */
int cvg_program_create(const char* integrand_src, const cvg_params* shape, const char* options,
                       cvg_program** out);
const char* cvg_program_log(const cvg_program* prog);
/* Size of the compiled code object (bytes), 0 if compilation failed. */
long long cvg_program_code_size(const cvg_program* prog);
/* Copy the gfx950 code object (up to cap bytes) for caching or inspection;
 * returns the bytes copied. */
long long cvg_program_get_code(const cvg_program* prog, void* buf, long long cap);
void cvg_program_destroy(cvg_program* prog);

/* Integrator state: allocates the device buffers and initialises the uniform
 * map and the initial allocation n_h = max(2, (int)(N / C)). */
int cvg_state_create(cvg_program* prog, const cvg_params* params, cvg_state** out);
void cvg_state_destroy(cvg_state* st);
/* Use this HIP stream (hipStream_t) for all device work; NULL = own stream. */
int cvg_state_set_stream(cvg_state* st, void* hip_stream);
/* Parameters of a parametrised integrand (host array of n_theta doubles). */
int cvg_state_set_theta(cvg_state* st, const double* theta, int n_theta);
/* Enqueue n_iter iterations.  use_graph: capture one iteration in a hipGraph
 * and replay it. */
int cvg_state_run(cvg_state* st, int n_iter, int use_graph);
/* Wait for the state's work and combine the iterations run so far. */
int cvg_state_result(cvg_state* st, cvg_result* out);
/* Per-iteration device timing of eager runs (use_graph = 0): HIP events
 * around the sampling kernel and around the whole iteration, on the state's
 * stream.  get_timing returns the number of timed iterations. */
int cvg_state_set_timing(cvg_state* st, int enable);
int cvg_state_get_timing(cvg_state* st, double* fill_ms, double* iter_ms, int cap);
/* Per-iteration estimates: I_k, sigma_k^2 and samples; returns the count. */
int cvg_state_history(cvg_state* st, double* I, double* sigma2, long long* nevals, int cap);
/* Per-iteration variance of I_k from the unbiased per-cube variances
 * (sum_h sigma_h^2 / (n_h - 1)); no reference counterpart. Returns the count. */
int cvg_state_history_unbiased(cvg_state* st, double* sigma2_unbiased, int cap);
/* Number of cubes of a state. */
long long cvg_state_n_cubes(const cvg_state* st);

/* State I/O (host buffers).  Map edges: n_dim * (n_intervals + 1). */
int cvg_state_get_map(cvg_state* st, double* x_edges);
int cvg_state_set_map(cvg_state* st, const double* x_edges);
/* Allocation n_h per cube (n_cubes ints). set_nh also rebuilds the sample layout;
 * every n_h must be >= 2, as the allocation rule guarantees. */
int cvg_state_get_nh(cvg_state* st, int* nh);
int cvg_state_set_nh(cvg_state* st, const int* nh);
/* d_h of the last iteration (n_cubes doubles). */
int cvg_state_get_dh(cvg_state* st, double* dh);
/* Per-cube sums of J f and (J f)^2 of the last iteration (n_cubes each). */
int cvg_state_get_cube_sums(cvg_state* st, double* jf, double* jf2);
/* Bin accumulators: weights (n_dim*n_intervals doubles), counts (uint32). */
int cvg_state_get_bins(cvg_state* st, double* weights, unsigned int* counts);
int cvg_state_set_bins(cvg_state* st, const double* weights, const unsigned int* counts);
/* Iteration counter (selects the Philox counter block). */
int cvg_state_set_iteration(cvg_state* st, int it);
/* Restore the per-iteration record of a saved run (no reference counterpart;
 * the reference keeps no state between calls): iterations [0, count) get
 * I_k, sigma_k^2, samples and the unbiased sigma_k^2 (NULL: sigma_k^2 again),
 * and both the device iteration counter and the
 * count of iterations done become count, so the next run continues with
 * iteration count and result() averages the restored iterations >= skip.
 * Together with set_map / set_nh this resumes or warm-starts a run. */
int cvg_state_set_history(cvg_state* st, const double* I, const double* sigma2, const long long* nevals,
                          const double* sigma2_unbiased, int count);

/* Single steps, for callers that drive their own loop. */
int cvg_state_fill(cvg_state* st);      /* one sampling pass with the current map */
int cvg_state_refine(cvg_state* st);    /* map refinement from the bin accumulators */
/* Allocation from given d_h (host, n_cubes) and D; rebuilds the layout. */
int cvg_state_reallocate(cvg_state* st, const double* dh, double D);

/* Samples [0, n) of iteration it under the current map and allocation:
 * cube index, bin ids (n*n_dim), point in the unit cube (n*n_dim), J f. */
int cvg_state_dump(cvg_state* st, int it, long long n, long long* cube, int* ids, double* x,
                   double* jf);

/* Differentiation of parametrised integrands (no reference counterpart).
 * The program must be compiled with the option "grad" and an integrand that
 * defines  vg_real integrand_grad(vg_real* x, const vg_real* theta, vg_real* g)
 * (returns f, writes df/dtheta).  set_record keeps the map and allocation of
 * every iteration >= skip; backward replays those frozen iterations (same
 * Philox counters), recomputing the samples, and returns with the forward's
 * inverse-variance weights held fixed:
 *   value = sum_k w_k I_k(theta) / sum_k w_k,  grad = d value / d theta,
 * their variances, for theta (n_theta host doubles) or the state's own
 * parameters if theta is NULL. */
int cvg_state_set_record(cvg_state* st, int enable);
int cvg_state_get_snapshot(cvg_state* st, int it, double* x_edges, int* nh);
int cvg_state_backward(cvg_state* st, const double* theta, double* grad, double* grad_var,
                       double* value, double* value_var);
/* cvg_state_backward without any allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as a
 * training step captured in a HIP graph or a serving loop kept off the
 * allocator.  Same arguments, kernels and results, bit for bit, as
 * cvg_state_backward; the replay buffers live in a static 64 MiB device array
 * of the library (one copy per GPU, held until the process ends) instead of
 * buffers the state allocates on its first backward.  Returns
 * CVG_ERR_CAPACITY, launching nothing, when the state's replay needs more
 * (about 8 (C + 1) + 8 (2 + 2 n_theta) (units + 1) bytes, units <= 65,536:
 * 41 MB for a million cubes and 32 parameters; up to about 4.5 M cubes).  Calls that overlap on one
 * GPU share the array and race; the replay itself does not time its kernels
 * (cvg_state_backward_timing keeps the last timed call's figure). */
int cvg_state_backward_noalloc(cvg_state* st, const double* theta, double* grad, double* grad_var,
                               double* value, double* value_var);
/* Device time (ms, HIP events on the state's stream) of the last backward
 * call's replay kernels (the per-iteration replay and its unit reduction),
 * summed over the recorded iterations. */
int cvg_state_backward_timing(cvg_state* st, double* replay_ms);

/* Batches (no reference counterpart): many independent integrals, each with
 * its own dimension, budget, strata, map and parameter block, advanced in
 * lockstep with one sampling launch per iteration.  The integrand has the form
 *   vg_real integrand(vg_real* x, const vg_real* theta, int d)
 * with theta pointing at the item's parameters.  Item i draws from Philox
 * stream i: it reproduces a single-integral run with stream_id = i. */
typedef struct cvg_batch cvg_batch;
typedef struct cvg_batch_item {
    int n_dim;            /* 1..max_dim */
    int n_strat;          /* 0: int((n_eval_it / 4)^(1/n_dim)) */
    long long n_eval_it;  /* samples per iteration */
    int theta_offset;     /* first parameter of this item: in [0, n_theta), or 0 when
                             n_theta is 0; the integrand reads only theta[offset ..
                             n_theta) (the device copy holds n_theta doubles, at
                             least one) */
} cvg_batch_item;
/* shape supplies n_intervals (<= 1024), precision and run_len; options may
 * hold block=N (threads per block), task_windows=N (windows per wave in one
 * task), fill_grid=0|1 (1, the default: one block per task; 0: a resident
 * grid striding over the tasks), bin_copies=1|2|4 (copies of the block's LDS
 * bins, lane l adding into copy l mod copies; default 2 when the blocks that
 * fill a CU's 16 waves still fit in LDS with them, else 1; CVG_ERR_INVALID
 * when they do not fit) and hipRTC options. */
int cvg_batch_program_create(const char* integrand_src, int max_dim, const cvg_params* shape,
                             const char* options, cvg_program** out);
/* common supplies n_intervals, precision, run_len, max_it, skip, alpha, beta, seed. */
int cvg_batch_create(cvg_program* prog, const cvg_params* common, const cvg_batch_item* items,
                     int n_items, const double* theta, int n_theta, cvg_batch** out);
void cvg_batch_destroy(cvg_batch* b);
int cvg_batch_run(cvg_batch* b, int n_iter, int use_graph);
/* Per item: weighted mean, error, chi2/dof, evaluations; loop_ms: device time;
 * error and chi2/dof from the unbiased per-cube variances (as in cvg_result).
 * Any output pointer may be NULL. */
int cvg_batch_results(cvg_batch* b, double* result, double* error, double* chi2_dof,
                      long long* nevals, double* loop_ms, double* error_unbiased,
                      double* chi2_dof_unbiased);
/* Per-iteration device timing of eager batch runs (use_graph = 0), as for a
 * single integral: HIP events around the one sampling launch of the batch and
 * around the whole iteration, on the batch's stream; get_timing returns the
 * number of timed iterations (no reference counterpart). */
int cvg_batch_set_timing(cvg_batch* b, int enable);
int cvg_batch_get_timing(cvg_batch* b, double* fill_ms, double* iter_ms, int cap);
int cvg_batch_history(cvg_batch* b, int item, double* I, double* sigma2, long long* nevals, int cap);
int cvg_batch_item_info(cvg_batch* b, int item, int* n_dim, int* n_strat, long long* n_cubes);
int cvg_batch_get_map(cvg_batch* b, int item, double* x_edges);
int cvg_batch_get_nh(cvg_batch* b, int item, int* nh);
int cvg_batch_get_cube_sums(cvg_batch* b, int item, double* jf, double* jf2);

/* One-shot call with the reference's argument list (cuda_wrapper.cu:290).
 * integrand_src replaces the numba PTX; options replaces the kernel PTX path
 * and may hold "fp32" (CVG_FP32C) and extra hipRTC options.  batch_size,
 * multi_cpu, cpu_num and gpu_thr are accepted and reported back unchanged;
 * the device path needs no sample batching or host-side cube mapping.
 * multi_gpu runs on the current device only.  The compiled program is kept
 * for later calls with the same source, shape, options and device (up to 16,
 * least recently used dropped; safe from several threads), so only the first
 * such call pays the hipRTC compile. */
int cvg_integrate_cuda(long long tot_eval, int max_it, int skip, long long max_batch_size,
                       int n_intervals, int n_strat, int n_dim, double alpha, double beta,
                       unsigned int seed, int multi_gpu, int multi_cpu, int gpu_num, int cpu_num,
                       double gpu_thr, const char* integrand_src, const char* options,
                       cvg_ref_result* out);

#ifdef __cplusplus
}
#endif

#endif /* CUVEGAS_H */
