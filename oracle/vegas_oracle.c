/*
 * vegas_oracle.c -- CPU restatement of the VEGAS+ iteration of
 * emiliantolo/cuvegas, driven by the same Philox4x32-10 stream as the GPU
 * build.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker or the timed CPU baseline.  The product (cuvegas_amd/) never
 * links, loads or calls it.
 *
 * Reference files followed (paths relative to the reference checkout):
 *   vegas_cuda/src/vegas_cuda_extension/cuda_kernel.cu:14-76  sampling kernel
 *   vegas_cuda/src/vegas_cuda_extension/cuda_wrapper.cu:26-207 auxiliary kernels
 *   vegas_cuda/src/vegas_cuda_extension/cuda_wrapper.cu:331-773 sizing, init, loop
 *   vegas_cuda/src/vegas_cuda_extension/python_wrapper.py:9-77 defaults, domain
 *
 * Where the reference is order-nondeterministic (CUB reductions, atomics), this
 * file fixes one order, and the GPU build implements the same order:
 *   - per-cube sums: sequential within a run of K consecutive samples,
 *     segmented scan across the 64 runs of a window in the steps of the GPU's
 *     DPP scan (within groups of 16 runs, then across groups), sequential
 *     fold across windows (orc_cube_sums);
 *   - reductions over cubes or bins: tree of 1024-wide strided pairwise sums,
 *     repeated over the partials (orc_tree_sum);
 *   - cumulative map weights: sequential inclusive scan.
 * Bin accumulators are plain sequential sums here; on the GPU they are
 * atomics, so they match to rounding only.
 *
 * Parity pins (see tests/test_oracle.py):
 *   orc_philox        Random123 known-answer vectors
 *   orc_log/exp/pow   libm within stated ulp bounds; special values
 *   orc_bin_of        exhaustive boundary words vs exact rational floor
 *   orc_tree_sum      exact on integer-valued data; fsum within bound
 *   orc_cube_sums     brute-force per-cube sums (exact on integer data)
 *   orc_results       hand-computed cases (constant f -> sigma^2 = 0)
 *   orc_alloc         invariants (n_h >= 2, uniform -> floor(N/C), sum bound)
 *   orc_refine        fixed point under uniform weights, monotone edges,
 *                     piecewise-constant weights -> analytic inverse CDF,
 *                     alpha = 0 -> uniform damping
 *   orc_run           closed-form integrals (pull < 4 sigma, chi2/dof < 2) of
 *                     every integrand it restates: polynomial, linear with a
 *                     domain, Gaussian (one and per-dimension centres), double
 *                     Gaussian, oscillatory, corner peak, mixture; f = 1 exact
 *   orc_grad_replay   value equal to the forward's I_k on the same samples;
 *                     gradient equal to central differences of the replayed
 *                     value; within 4 sigma of the erf closed-form gradient
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))
#define ORC_WAVE 64

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11).                                     */
/* ------------------------------------------------------------------------ */

ORC_EXPORT void orc_philox(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------------ */
/* Deterministic log / exp / pow (same method as the device routine):        */
/* atanh series for log on m in [sqrt(1/2), sqrt(2)), Cody-Waite reduction   */
/* plus Taylor series for exp.                                               */
/* ------------------------------------------------------------------------ */

static const double LN2_HI = 6.93147180369123816490e-01;
static const double LN2_LO = 1.90821492927058770002e-10;
static const double INV_LN2 = 1.44269504088896338700e+00;
static const double SQRT2 = 1.41421356237309514547e+00;

static double from_bits(uint64_t b) { double x; memcpy(&x, &b, 8); return x; }
static uint64_t to_bits(double x) { uint64_t b; memcpy(&b, &x, 8); return b; }

ORC_EXPORT double orc_log(double x) {
    if (x != x) return x;
    if (x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (isinf(x)) return x;
    int adj = 0;
    if (x < 2.2250738585072014e-308) { x *= 18014398509481984.0; adj = -54; }
    uint64_t b = to_bits(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023 + adj;
    double m = from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > SQRT2) { m *= 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double p = 1.0 / 23.0;
    p = fma(p, z, 1.0 / 21.0);
    p = fma(p, z, 1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    double logm = fma((s + s) * z, p, s + s);
    double de = (double)e;
    return fma(de, LN2_HI, fma(de, LN2_LO, logm));
}

ORC_EXPORT double orc_exp(double y) {
    if (y != y) return y;
    double yc = y < -746.0 ? -746.0 : (y > 710.0 ? 710.0 : y);  /* 2^k then under/overflows in ldexp */
    double kd = floor(fma(yc, INV_LN2, 0.5));
    double r = fma(-kd, LN2_LO, fma(-kd, LN2_HI, yc));
    /* Taylor polynomial sum_{i<=13} r^i / i!: r^4..r^13 tail by Estrin, leading terms by Horner (as vg_exp) */
    double r2 = r * r, r4 = r2 * r2;
    double q0 = fma(r, 1.0 / 120.0, 1.0 / 24.0);
    double q1 = fma(r, 1.0 / 5040.0, 1.0 / 720.0);
    double q2 = fma(r, 1.0 / 362880.0, 1.0 / 40320.0);
    double q3 = fma(r, 1.0 / 39916800.0, 1.0 / 3628800.0);
    double q4 = fma(r, 1.0 / 6227020800.0, 1.0 / 479001600.0);
    double s0 = fma(q1, r2, q0);
    double s1 = fma(q3, r2, q2);
    double t = fma(fma(q4, r4, s1), r4, s0);
    double p = fma(t, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)kd);
}

ORC_EXPORT double orc_pow(double x, double a) {
    if (x != x || a != a || x < 0.0) return NAN;
    if (a == 0.0 || x == 1.0) return 1.0;
    if (x == 0.0) return a > 0.0 ? 0.0 : INFINITY;
    if (isinf(x)) return a > 0.0 ? INFINITY : 0.0;
    return orc_exp(a * orc_log(x));
}

/* (int) conversion as a truncating hardware convert: NaN -> 0, saturating. */
ORC_EXPORT int orc_sat_int(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return INT32_MIN;
    return (int)v;
}

/* ------------------------------------------------------------------------ */
/* Sample -> bin, exactly in integers (u = (2r+1)/2^33).                      */
/* Reference: y = u*dy + k*dy; id = (int)(y*n); x = x_id + dx_id*(y*n - id)  */
/* (cuda_kernel.cu:35-46).                                                   */
/* ------------------------------------------------------------------------ */

ORC_EXPORT void orc_bin_of(uint32_t r, uint32_t kn, uint32_t nbins, uint32_t nstrat,
                           uint32_t* id, uint32_t* rem, uint64_t* low) {
    uint64_t L = ((uint64_t)r * 2u + 1u) * (uint64_t)nbins;
    uint32_t Q = kn + (uint32_t)(L >> 33);
    *id = Q / nstrat;
    *rem = Q - *id * nstrat;
    *low = L & ((1ull << 33) - 1ull);
}

/* ------------------------------------------------------------------------ */
/* Fixed-order reductions.                                                   */
/* ------------------------------------------------------------------------ */

/* One level: chunks of 1024 reduced by strided pairwise sums
 * (t[i] += t[i+512], then +256, ... , +1), missing entries are 0.0. */
static long long tree_level(const double* a, long long m, double* out) {
    long long nchunks = (m + 1023) / 1024;
    double t[1024];
    for (long long c = 0; c < nchunks; ++c) {
        for (int i = 0; i < 1024; ++i) {
            long long j = c * 1024 + i;
            t[i] = j < m ? a[j] : 0.0;
        }
        for (int stride = 512; stride >= 1; stride >>= 1)
            for (int i = 0; i < stride; ++i) t[i] = t[i] + t[i + stride];
        out[c] = t[0];
    }
    return nchunks;
}

/* Sum of a[0..m): one level always, then levels over the partials until one
 * value remains.  Replaces cub::DeviceReduce::Sum (cuda_wrapper.cu:721,726,
 * 729,733-734). */
ORC_EXPORT double orc_tree_sum(const double* a, long long m) {
    if (m <= 0) return 0.0;
    long long nc = (m + 1023) / 1024;
    double* buf = (double*)malloc(sizeof(double) * (size_t)nc);
    long long cur = tree_level(a, m, buf);
    while (cur > 1) cur = tree_level(buf, cur, buf);
    double r = buf[0];
    free(buf);
    return r;
}

/* ------------------------------------------------------------------------ */
/* Per-cube sums in the fixed order (replaces the per-sample atomicAdd on     */
/* JF/JF2, cuda_kernel.cu:67-70).  The sample stream s = 0..S-1 (cube-major) */
/* is cut into runs of K samples; 64 consecutive runs form a window.          */
/*  level 0: within a run, consecutive samples of one cube are summed in     */
/*           order starting from 0.0 ("pieces");                             */
/*  level 1: within a window, each run's last piece enters a segmented       */
/*           inclusive scan over the 64 runs in the GPU's wave order: steps  */
/*           of 1, 2, 4, 8 runs inside each group of 16 runs, then run 15 of */
/*           the group before into groups 1 and 3, then run 31 into groups 2 */
/*           and 3 (new = left + right when the keys match); a cube ending   */
/*           inside a run as that run's first piece takes scan(previous run) */
/*           + piece;                                                        */
/*  level 2: a cube spanning windows folds its window partials left to right.*/
/* The GPU fill kernel implements the same three levels.                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    long long S, C, K;
    const long long* nh_sum;   /* C+1 entries, exclusive prefix, nh_sum[C] = S */
    double *JF, *JF2;          /* C */
    double *wf1, *wf2, *wl1, *wl2; /* per window: first-cube and last-cube partials */
    /* scratch for one window */
    double f1[ORC_WAVE], f2[ORC_WAVE], l1[ORC_WAVE], l2[ORC_WAVE];
    long long hF[ORC_WAVE], hL[ORC_WAVE], startF[ORC_WAVE], startL[ORC_WAVE];
    int hasF[ORC_WAVE], cont[ORC_WAVE], active[ORC_WAVE];
} cube_acc;

static void window_finish(cube_acc* A, long long w) {
    double s1[ORC_WAVE], s2[ORC_WAVE], n1[ORC_WAVE], n2[ORC_WAVE];
    long long key[ORC_WAVE];
    for (int l = 0; l < ORC_WAVE; ++l) {
        key[l] = A->active[l] ? A->hL[l] : -1;
        s1[l] = A->active[l] ? A->l1[l] : 0.0;
        s2[l] = A->active[l] ? A->l2[l] : 0.0;
    }
    /* six steps; step t takes lane src(t, l) (or none, -1) into lane l */
    for (int t = 0; t < 6; ++t) {
        for (int l = 0; l < ORC_WAVE; ++l) {
            const int row = l >> 4;
            int src = -1;
            if (t < 4) {
                const int off = 1 << t;                       /* row_shr 1, 2, 4, 8 */
                if ((l & 15) >= off) src = l - off;
            } else if (t == 4) {
                if (row == 1 || row == 3) src = 16 * row - 1; /* row_bcast 15 */
            } else {
                if (row >= 2) src = 31;                       /* row_bcast 31 */
            }
            if (src >= 0 && key[src] == key[l]) {
                n1[l] = s1[src] + s1[l];
                n2[l] = s2[src] + s2[l];
            } else {
                n1[l] = s1[l];
                n2[l] = s2[l];
            }
        }
        memcpy(s1, n1, sizeof s1);
        memcpy(s2, n2, sizeof s2);
    }
    const long long wstart = w * ORC_WAVE * A->K;
    for (int l = 0; l < ORC_WAVE; ++l) {
        if (!A->active[l]) continue;
        if (A->hasF[l]) {
            double t1 = A->f1[l], t2 = A->f2[l];
            if (l > 0) { t1 = s1[l - 1] + t1; t2 = s2[l - 1] + t2; }
            if (A->startF[l] < wstart) { A->wf1[w] = t1; A->wf2[w] = t2; }
            else { A->JF[A->hF[l]] = t1; A->JF2[A->hF[l]] = t2; }
        }
        if (!A->cont[l]) {
            if (A->startL[l] < wstart) { A->wf1[w] = s1[l]; A->wf2[w] = s2[l]; }
            else { A->JF[A->hL[l]] = s1[l]; A->JF2[A->hL[l]] = s2[l]; }
        } else if (l == ORC_WAVE - 1) {
            A->wl1[w] = s1[l]; A->wl2[w] = s2[l];
            if (A->startL[l] < wstart) { A->wf1[w] = s1[l]; A->wf2[w] = s2[l]; }
        }
    }
}

/* Level 2 for every cube spanning windows wa..wb (m = wb - wa + 1 > 1 window
 * partials: win_last[wa], win_first[wa+1..wb]): a left fold when m <= 64
 * (done inside the GPU results kernel), else the 1024-wide tree of
 * orc_tree_sum (a separate GPU kernel, one block per such cube). */
static void cube_fold(cube_acc* A) {
    const long long WS = ORC_WAVE * A->K;
    for (long long h = 0; h < A->C; ++h) {
        long long st = A->nh_sum[h], en = A->nh_sum[h + 1];
        if (en <= st) continue;
        long long wa = st / WS, wb = (en - 1) / WS;
        if (wa == wb) continue;
        const long long m = wb - wa + 1;
        if (m <= 64) {
            double t1 = A->wl1[wa], t2 = A->wl2[wa];
            for (long long w = wa + 1; w <= wb; ++w) { t1 = t1 + A->wf1[w]; t2 = t2 + A->wf2[w]; }
            A->JF[h] = t1; A->JF2[h] = t2;
        } else {
            double* v1 = (double*)malloc(sizeof(double) * (size_t)m);
            double* v2 = (double*)malloc(sizeof(double) * (size_t)m);
            v1[0] = A->wl1[wa]; v2[0] = A->wl2[wa];
            for (long long k = 1; k < m; ++k) { v1[k] = A->wf1[wa + k]; v2[k] = A->wf2[wa + k]; }
            A->JF[h] = orc_tree_sum(v1, m);
            A->JF2[h] = orc_tree_sum(v2, m);
            free(v1); free(v2);
        }
    }
}

/* Sample evaluation callback: returns jf and jf^2 of global sample s in cube h
 * and accumulates the bin histograms. */
typedef void (*sample_fn)(void* ctx, long long s, long long h, double* jf, double* jf2);

/* Windows [w0, w1) of the sample stream (levels 0 and 1). */
static void cube_sums_range(cube_acc* A, long long w0, long long w1, sample_fn fn, void* ctx) {
    const long long S = A->S, K = A->K;
    long long h = 0;   /* cube of the current sample, advances monotonically */
    if (w0 > 0) {      /* largest h with nh_sum[h] <= first sample of the range */
        long long s0 = w0 * ORC_WAVE * K, lo = 0, hi = A->C;
        while (hi - lo > 1) { long long m = (lo + hi) / 2; if (A->nh_sum[m] <= s0) lo = m; else hi = m; }
        h = lo;
    }
    for (long long w = w0; w < w1; ++w) {
        for (int l = 0; l < ORC_WAVE; ++l) {
            long long r = w * ORC_WAVE + l, rs = r * K;
            long long re = rs + K < S ? rs + K : S;
            A->active[l] = rs < S;
            A->hasF[l] = 0;
            if (!A->active[l]) continue;
            while (A->nh_sum[h + 1] <= rs) ++h;
            long long cstart = A->nh_sum[h], cend = A->nh_sum[h + 1];
            int sb = cstart < rs, npieces = 0;
            double a1 = 0.0, a2 = 0.0;
            for (long long s = rs; s < re; ++s) {
                if (s == cend) {
                    if (npieces == 0 && sb) {
                        A->hasF[l] = 1; A->hF[l] = h; A->startF[l] = cstart;
                        A->f1[l] = a1; A->f2[l] = a2;
                    } else {
                        A->JF[h] = a1; A->JF2[h] = a2;
                    }
                    ++npieces;
                    ++h;
                    while (A->nh_sum[h + 1] <= s) ++h;  /* (nh >= 1 so no skip) */
                    cstart = cend; cend = A->nh_sum[h + 1];
                    a1 = 0.0; a2 = 0.0;
                }
                double v1, v2;
                fn(ctx, s, h, &v1, &v2);
                a1 = a1 + v1; a2 = a2 + v2;
            }
            A->hL[l] = h; A->startL[l] = cstart;
            A->l1[l] = a1; A->l2[l] = a2;
            A->cont[l] = cend > re;
        }
        window_finish(A, w);
    }
}

static void cube_sums_driver(cube_acc* A, sample_fn fn, void* ctx) {
    const long long WS = ORC_WAVE * A->K;
    cube_sums_range(A, 0, (A->S + WS - 1) / WS, fn, ctx);
    cube_fold(A);
}

typedef struct { const double *v1, *v2; } vals_ctx;
static void vals_fn(void* c, long long s, long long h, double* a, double* b) {
    (void)h;
    const vals_ctx* v = (const vals_ctx*)c;
    *a = v->v1[s]; *b = v->v2[s];
}

static int cube_acc_init(cube_acc* A, long long C, const long long* nh_sum, long long K,
                         double* JF, double* JF2) {
    memset(A, 0, sizeof *A);
    A->C = C; A->S = nh_sum[C]; A->K = K; A->nh_sum = nh_sum; A->JF = JF; A->JF2 = JF2;
    long long nwin = (A->S + ORC_WAVE * K - 1) / (ORC_WAVE * K) + 1;
    A->wf1 = (double*)calloc((size_t)nwin, 8); A->wf2 = (double*)calloc((size_t)nwin, 8);
    A->wl1 = (double*)calloc((size_t)nwin, 8); A->wl2 = (double*)calloc((size_t)nwin, 8);
    return (A->wf1 && A->wf2 && A->wl1 && A->wl2) ? 0 : -1;
}
static void cube_acc_free(cube_acc* A) { free(A->wf1); free(A->wf2); free(A->wl1); free(A->wl2); }

/* Per-cube sums of given per-sample values (testing entry point). */
ORC_EXPORT int orc_cube_sums(long long C, const long long* nh_sum, long long K,
                             const double* v1, const double* v2, double* JF, double* JF2) {
    cube_acc A;
    if (cube_acc_init(&A, C, nh_sum, K, JF, JF2)) return -1;
    vals_ctx vc = {v1, v2};
    cube_sums_driver(&A, vals_fn, &vc);
    cube_acc_free(&A);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Per-cube results (cuda_wrapper.cu:26-49), same expression order.          */
/* sums[0] = sum I_h, sums[1] = sum sigma_h^2/n_h, sums[2] = D = sum d_h,    */
/* sums[3] = sum sigma_h^2/(n_h - 1) (the unbiased per-cube variance, an     */
/* addition reported beside the reference's).                                */
/* ------------------------------------------------------------------------ */

ORC_EXPORT void orc_results(long long C, double V, double beta, const int* nh,
                            const double* JF, const double* JF2, double* dh, double* sums) {
    double* sr = (double*)malloc(sizeof(double) * (size_t)C);
    double* ss = (double*)malloc(sizeof(double) * (size_t)C);
    double* su = (double*)malloc(sizeof(double) * (size_t)C);
    for (long long i = 0; i < C; ++i) {
        int neval = nh[i];
        double Ih = JF[i] / neval * V;
        double Sig2 = (JF2[i] / neval * V * V) - (Ih * Ih);
        sr[i] = Ih;
        ss[i] = Sig2 / neval;
        su[i] = Sig2 / (neval - 1);
        double u = V * JF[i] / neval;
        double t = V * V * JF2[i] / neval - u * u;
        if (t < 0) t = 0;
        dh[i] = orc_pow(t, beta);
    }
    sums[0] = orc_tree_sum(sr, C);
    sums[1] = orc_tree_sum(ss, C);
    sums[2] = orc_tree_sum(dh, C);
    sums[3] = orc_tree_sum(su, C);
    free(sr); free(ss); free(su);
}

/* ------------------------------------------------------------------------ */
/* Stratification reallocation (cuda_wrapper.cu:118-132) and the exclusive  */
/* prefix sum that places each cube's samples (cuda_wrapper.cu:634).         */
/* ------------------------------------------------------------------------ */

ORC_EXPORT void orc_alloc(long long C, const double* dh, double D, long long N, int* nh) {
    for (long long i = 0; i < C; ++i) {
        double f = dh[i] / D;
        int v = orc_sat_int(f * (double)N);
        nh[i] = v < 2 ? 2 : v;
    }
}

ORC_EXPORT long long orc_prefix(long long C, const int* nh, long long* nh_sum) {
    long long acc = 0;
    for (long long i = 0; i < C; ++i) { nh_sum[i] = acc; acc += nh[i]; }
    nh_sum[C] = acc;
    return acc;
}

/* ------------------------------------------------------------------------ */
/* Map refinement (cuda_wrapper.cu:51-90,134-191,719-746) for one dimension: */
/* normalise, smooth, alpha-damp, cumulate, invert.  x_in/x_out: n+1 edges.  */
/* If the bin weights sum to zero (or NaN) the map is left unchanged (the    */
/* reference would divide by zero).                                          */
/* ------------------------------------------------------------------------ */

ORC_EXPORT void orc_refine_dim(int n, double alpha, const double* weights, const uint32_t* counts,
                               const double* x_in, double* x_out) {
    double* w = (double*)calloc((size_t)n, sizeof(double));
    double* d = (double*)malloc(sizeof(double) * (size_t)n);
    double* cum = (double*)malloc(sizeof(double) * (size_t)n);
    for (int k = 0; k < n; ++k) w[k] = counts[k] != 0 ? weights[k] / (double)counts[k] : weights[k];
    double dsum = orc_tree_sum(w, n);
    memcpy(x_out, x_in, sizeof(double) * (size_t)(n + 1));
    if (dsum > 0.0) {
        for (int k = 0; k < n; ++k) {
            double t;
            if (k == 0) t = (7.0 * w[k] + w[k + 1]) / (8.0 * dsum);
            else if (k == n - 1) t = (w[k - 1] + 7.0 * w[k]) / (8.0 * dsum);
            else t = (w[k - 1] + 6.0 * w[k] + w[k + 1]) / (8.0 * dsum);
            if (t != 0) t = orc_pow((t - 1.0) / orc_log(t), alpha);
            d[k] = t;
        }
        double S = orc_tree_sum(d, n);
        double acc = 0.0;
        for (int k = 0; k < n; ++k) { acc = acc + d[k]; cum[k] = acc; }
        double delta = S / n;
        for (int j = 0; j < n - 1; ++j) {
            int target = j + 1, lo = 0, hi = n;   /* first k with (int)(cum_k/delta) >= target */
            while (lo < hi) {
                int mid = (lo + hi) / 2;
                if (orc_sat_int(cum[mid] / delta) >= target) hi = mid; else lo = mid + 1;
            }
            int k = lo < n ? lo : n - 1;
            double a = target * delta - (k > 0 ? cum[k - 1] : 0.0);
            x_out[j + 1] = x_in[k] + a / d[k] * (x_in[k + 1] - x_in[k]);
        }
    }
    free(w); free(d); free(cum);
}

/* ------------------------------------------------------------------------ */
/* Integrands (C restatements of the sources in cuvegas_amd/integrands.py;   */
/* ORC_F_GAUSSC and the oscillatory and corner forms also restate the        */
/* families of the batched source, cuvegas_amd/batch.py FAMILY_SOURCE).      */
/* ------------------------------------------------------------------------ */

enum {
    ORC_F_ONE = 0, ORC_F_LINEAR = 1, ORC_F_POLY = 2, ORC_F_GAUSS = 3, ORC_F_DGAUSS = 4,
    ORC_F_OSC = 5, ORC_F_CORNER = 6, ORC_F_MIXTURE = 7, ORC_F_GAUSSC = 8
};

static const double ORC_PI = 3.14159265358979323846;

/* 4-component Gaussian mixture, theta = {w[4], s[4], mu[4][d]}; writes
 * df/dtheta to g when g != NULL. */
static double mix64(const double* x, int d, const double* th, double* g) {
    const double c = sqrt(2.0 * ORC_PI);
    double f = 0.0;
    for (int k = 0; k < 4; ++k) {
        double w = th[k], s = th[4 + k], q = 0.0;
        const double* mu = th + 8 + k * d;
        double is2 = 1.0 / (s * s), is = 1.0 / s, m = -0.5 * is2;
        double norm = 1.0;
        for (int i = 0; i < d; ++i) norm = norm * (s * c);
        double inorm = 1.0 / norm;
        for (int i = 0; i < d; ++i) { double a = x[i] - mu[i]; q = fma(a, a, q); }
        double pk = orc_exp(q * m) * inorm;
        f = f + w * pk;
        if (g) {
            double wp = w * pk;
            g[k] = pk;
            g[4 + k] = wp * (q * is2 - (double)d) * is;
            for (int i = 0; i < d; ++i) g[8 + k * d + i] = wp * (x[i] - mu[i]) * is2;
        }
    }
    return f;
}

static float mix32(const float* x, int d, const float* th, float* g) {
    const float c = sqrtf(2.0f * (float)ORC_PI);
    float f = 0.0f;
    for (int k = 0; k < 4; ++k) {
        float w = th[k], s = th[4 + k], q = 0.0f;
        const float* mu = th + 8 + k * d;
        float is2 = 1.0f / (s * s), is = 1.0f / s, m = -0.5f * is2;
        float norm = 1.0f;
        for (int i = 0; i < d; ++i) norm = norm * (s * c);
        float inorm = 1.0f / norm;
        for (int i = 0; i < d; ++i) { float a = x[i] - mu[i]; q = fmaf(a, a, q); }
        float pk = expf(q * m) * inorm;
        f = f + w * pk;
        if (g) {
            float wp = w * pk;
            g[k] = pk;
            g[4 + k] = wp * (q * is2 - (float)d) * is;
            for (int i = 0; i < d; ++i) g[8 + k * d + i] = wp * (x[i] - mu[i]) * is2;
        }
    }
    return f;
}

static double f64_eval(int id, const double* x, int d, const double* p) {
    switch (id) {
    case ORC_F_ONE: return 1.0;
    case ORC_F_LINEAR: { double f = 0.0; for (int i = 0; i < d; ++i) f = f + x[i]; return f; }
    case ORC_F_POLY: { double f = 1.0; for (int i = 0; i < d; ++i) f = f * (6.0 * x[i] * (1.0 - x[i])); return f; }
    case ORC_F_GAUSS: { /* p = {mu, sigma, norm, m}: norm * exp(m * sum (x-mu)^2), m = -1/(2 sigma^2) */
        double q = 0.0;
        for (int i = 0; i < d; ++i) { double a = x[i] - p[0]; q = fma(a, a, q); }
        return p[2] * orc_exp(q * p[3]);
    }
    case ORC_F_DGAUSS: { /* p = {c1, c2, sigma, norm, m} */
        double q1 = 0.0, q2 = 0.0;
        for (int i = 0; i < d; ++i) {
            double a = x[i] - p[0], b = x[i] - p[1];
            q1 = fma(a, a, q1); q2 = fma(b, b, q2);
        }
        return 0.5 * p[3] * (orc_exp(q1 * p[4]) + orc_exp(q2 * p[4]));
    }
    case ORC_F_OSC: { /* p = {u, a_0..a_{d-1}} */
        double t = 2.0 * ORC_PI * p[0];
        for (int i = 0; i < d; ++i) t = t + p[1 + i] * x[i];
        return cos(t);
    }
    case ORC_F_CORNER: { /* p = {a_0..a_{d-1}} */
        double t = 1.0;
        for (int i = 0; i < d; ++i) t = t + p[i] * x[i];
        double r = 1.0 / t, f = 1.0;
        for (int i = 0; i <= d; ++i) f = f * r;
        return f;
    }
    case ORC_F_MIXTURE: return mix64(x, d, p, NULL);
    case ORC_F_GAUSSC: { /* p = {norm, m, c_0..c_{d-1}}: norm * exp(m * sum (x_i - c_i)^2) */
        double q = 0.0;
        for (int i = 0; i < d; ++i) { double a = x[i] - p[2 + i]; q = fma(a, a, q); }
        return p[0] * orc_exp(q * p[1]);
    }
    }
    return NAN;
}

static float f32_eval(int id, const float* x, int d, const float* p) {
    switch (id) {
    case ORC_F_ONE: return 1.0f;
    case ORC_F_LINEAR: { float f = 0.0f; for (int i = 0; i < d; ++i) f = f + x[i]; return f; }
    case ORC_F_POLY: { float f = 1.0f; for (int i = 0; i < d; ++i) f = f * (6.0f * x[i] * (1.0f - x[i])); return f; }
    case ORC_F_GAUSS: {
        float q = 0.0f;
        for (int i = 0; i < d; ++i) { float a = x[i] - p[0]; q = fmaf(a, a, q); }
        return p[2] * expf(q * p[3]);
    }
    case ORC_F_DGAUSS: {
        float q1 = 0.0f, q2 = 0.0f;
        for (int i = 0; i < d; ++i) {
            float a = x[i] - p[0], b = x[i] - p[1];
            q1 = fmaf(a, a, q1); q2 = fmaf(b, b, q2);
        }
        return 0.5f * p[3] * (expf(q1 * p[4]) + expf(q2 * p[4]));
    }
    case ORC_F_OSC: {
        float t = 2.0f * (float)ORC_PI * p[0];
        for (int i = 0; i < d; ++i) t = t + p[1 + i] * x[i];
        return cosf(t);
    }
    case ORC_F_CORNER: {
        float t = 1.0f;
        for (int i = 0; i < d; ++i) t = t + p[i] * x[i];
        float r = 1.0f / t, f = 1.0f;
        for (int i = 0; i <= d; ++i) f = f * r;
        return f;
    }
    case ORC_F_MIXTURE: return mix32(x, d, p, NULL);
    case ORC_F_GAUSSC: {
        float q = 0.0f;
        for (int i = 0; i < d; ++i) { float a = x[i] - p[2 + i]; q = fmaf(a, a, q); }
        return p[0] * expf(q * p[1]);
    }
    }
    return NAN;
}

/* ------------------------------------------------------------------------ */
/* Full integrator (cuda_wrapper.cu:290-872).                                */
/* ------------------------------------------------------------------------ */

typedef struct {
    int d, nbins, nstrat, max_it, skip, K;
    long long n_eval_it;
    double alpha, beta;
    uint32_t seed, stream;
    int precision;            /* 0 = fp64, 1 = fp32 sampling/eval */
    int integrand;            /* ORC_F_* */
    const double* iparams;    /* integrand parameters */
    int n_iparams;
    const double* domain;     /* NULL or 2*d: start[d], range[d] */
    double domain_vol;
    int iters;                /* iterations to run (<= max_it) */
    int threads;              /* OpenMP threads for the sampling pass (<= 1: serial) */
} orc_params;

typedef struct {
    double* hist_I;           /* iters: per-iteration estimate */
    double* hist_S2;          /* iters: per-iteration variance */
    long long* hist_n;        /* iters: samples used */
    double* x_edges;          /* d*(n+1): final map */
    int* nh;                  /* C: final allocation */
    int* nh_first;            /* C or NULL: allocation after iteration 0 */
    double* x_first;          /* d*(n+1) or NULL: map after iteration 0 */
    double* dh_first;         /* C or NULL: d_h of iteration 0 */
    double* JF_first;         /* C or NULL: per-cube sums of iteration 0 */
    double* JF2_first;
    double* w_first;          /* d*n or NULL: bin weights of iteration 0 */
    uint32_t* c_first;        /* d*n or NULL: bin counts of iteration 0 */
    double result, error, chi2_dof;
    long long tot_nevals;
} orc_out;

typedef struct {
    const orc_params* P;
    long long C;
    int it;
    const double* x_edges;    /* d*(n+1) */
    const double* dx;         /* d*n */
    const float* ef;          /* d*(n+1) */
    const float* dxf;         /* d*n */
    double* weights; uint32_t* counts;
    float fparams[512];
    /* optional per-sample dump */
    long long dump_n;
    int* dump_ids; double* dump_x; double* dump_jf; long long* dump_cube;
} fill_ctx;

static void sample_fn_impl(void* c, long long s, long long h, double* out_jf, double* out_jf2) {
    fill_ctx* F = (fill_ctx*)c;
    const orc_params* P = F->P;
    const int d = P->d, n = P->nbins, E = n + 1;
    const uint32_t ns = (uint32_t)P->nstrat;
    uint32_t words[68];
    for (int g = 0; g < (d + 3) / 4; ++g) {
        uint32_t ctr[4] = {(uint32_t)s, (uint32_t)((uint64_t)s >> 32), (uint32_t)F->it, (uint32_t)g};
        orc_philox(ctr, P->seed, P->stream, &words[4 * g]);
    }
    uint32_t ids[64];
    long long tmp = h;
    double jf, jf2;
    if (P->precision == 0) {
        double x[64], jac = 1.0;
        const double inv_ns = 1.0 / (double)ns;
        for (int i = 0; i < d; ++i) {
            uint32_t k = (uint32_t)(tmp % ns); tmp /= ns;
            uint32_t id, rem; uint64_t low;
            orc_bin_of(words[i], k * (uint32_t)n, (uint32_t)n, ns, &id, &rem, &low);
            double frac = ((double)rem + (double)low * 0x1p-33) * inv_ns;
            x[i] = fma(F->dx[i * n + id], frac, F->x_edges[i * E + id]);
            ids[i] = id;
        }
        for (int i = 0; i < d; ++i) jac = jac * ((double)n * F->dx[i * n + ids[i]]);
        if (F->dump_x && s < F->dump_n) for (int i = 0; i < d; ++i) F->dump_x[s * d + i] = x[i];
        if (P->domain) for (int i = 0; i < d; ++i) x[i] = x[i] * P->domain[d + i] + P->domain[i];
        double f = f64_eval(P->integrand, x, d, P->iparams);
        if (P->domain) f = P->domain_vol * f;
        jf = f * jac;
        jf2 = jf * jf;
    } else {
        float x[64], jac = 1.0f;
        const float inv_ns = 1.0f / (float)ns;
        for (int i = 0; i < d; ++i) {
            uint32_t k = (uint32_t)(tmp % ns); tmp /= ns;
            uint32_t id, rem; uint64_t low;
            orc_bin_of(words[i], k * (uint32_t)n, (uint32_t)n, ns, &id, &rem, &low);
            float frac = ((float)rem + (float)(uint32_t)(low >> 9) * 0x1p-24f) * inv_ns;
            x[i] = fmaf(F->dxf[i * n + id], frac, F->ef[i * E + id]);
            ids[i] = id;
        }
        for (int i = 0; i < d; ++i) jac = jac * ((float)n * F->dxf[i * n + ids[i]]);
        if (F->dump_x && s < F->dump_n) for (int i = 0; i < d; ++i) F->dump_x[s * d + i] = (double)x[i];
        if (P->domain) for (int i = 0; i < d; ++i) x[i] = x[i] * (float)P->domain[d + i] + (float)P->domain[i];
        float f = f32_eval(P->integrand, x, d, F->fparams);
        if (P->domain) f = (float)P->domain_vol * f;
        float jff = f * jac;
        jf = (double)jff;
        jf2 = jf * jf;
    }
    for (int i = 0; i < d; ++i) {
        F->weights[i * n + ids[i]] += jf2;
        F->counts[i * n + ids[i]] += 1;
    }
    if (F->dump_ids && s < F->dump_n) {
        for (int i = 0; i < d; ++i) F->dump_ids[s * d + i] = (int)ids[i];
        F->dump_jf[s] = jf;
        F->dump_cube[s] = h;
    }
    *out_jf = jf; *out_jf2 = jf2;
}

static int ipow_checked(int b, int e, long long* out) {
    long long r = 1;
    for (int i = 0; i < e; ++i) { r *= b; if (r > 2147483647LL) return -1; }
    *out = r;
    return 0;
}

static void build_f32_map(int d, int n, const double* x_edges, float* ef, float* dxf) {
    for (int i = 0; i < d; ++i) {
        for (int k = 0; k <= n; ++k) ef[i * (n + 1) + k] = (float)x_edges[i * (n + 1) + k];
        for (int k = 0; k < n; ++k) dxf[i * n + k] = ef[i * (n + 1) + k + 1] - ef[i * (n + 1) + k];
    }
}

ORC_EXPORT int orc_run(const orc_params* P, orc_out* O) {
    const int d = P->d, n = P->nbins, E = n + 1;
    long long C;
    if (d < 1 || d > 64 || n < 2 || P->nstrat < 1 || P->K < 1) return -1;
    if (ipow_checked(P->nstrat, d, &C)) return -2;
    const double V = pow(1.0 / P->nstrat, d);
    const long long N = P->n_eval_it;

    double* x_edges = (double*)malloc(sizeof(double) * (size_t)(d * E));
    double* x_new = (double*)malloc(sizeof(double) * (size_t)E);
    double* dx = (double*)malloc(sizeof(double) * (size_t)(d * n));
    float* ef = (float*)malloc(sizeof(float) * (size_t)(d * E));
    float* dxf = (float*)malloc(sizeof(float) * (size_t)(d * n));
    double* weights = (double*)malloc(sizeof(double) * (size_t)(d * n));
    uint32_t* counts = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(d * n));
    int* nh = (int*)malloc(sizeof(int) * (size_t)C);
    long long* nh_sum = (long long*)malloc(sizeof(long long) * (size_t)(C + 1));
    double* JF = (double*)malloc(sizeof(double) * (size_t)C);
    double* JF2 = (double*)malloc(sizeof(double) * (size_t)C);
    double* dh = (double*)malloc(sizeof(double) * (size_t)C);
    if (!x_edges || !x_new || !dx || !ef || !dxf || !weights || !counts || !nh || !nh_sum || !JF || !JF2 || !dh)
        return -3;

    /* uniform map and initial allocation (cuda_wrapper.cu:464-480) */
    const double step = 1.0 / n;
    for (int i = 0; i < d; ++i) {
        for (int j = 0; j < E; ++j) x_edges[i * E + j] = j * step;
        for (int j = 0; j < n; ++j) dx[i * n + j] = x_edges[i * E + j + 1] - x_edges[i * E + j];
    }
    int nh0 = orc_sat_int(1.0 / (double)C * (double)N);
    if (nh0 < 2) nh0 = 2;
    for (long long h = 0; h < C; ++h) nh[h] = nh0;

    fill_ctx F;
    memset(&F, 0, sizeof F);
    F.P = P; F.C = C; F.x_edges = x_edges; F.dx = dx; F.ef = ef; F.dxf = dxf;
    F.weights = weights; F.counts = counts;
    for (int i = 0; i < P->n_iparams && i < 512; ++i) F.fparams[i] = (float)P->iparams[i];

    O->tot_nevals = 0;
    int rc = 0;
    for (int it = 0; it < P->iters; ++it) {
        long long S = orc_prefix(C, nh, nh_sum);
        O->tot_nevals += S;
        memset(weights, 0, sizeof(double) * (size_t)(d * n));
        memset(counts, 0, sizeof(uint32_t) * (size_t)(d * n));
        if (P->precision == 1) build_f32_map(d, n, x_edges, ef, dxf);
        F.it = it;
        cube_acc A;
        if (cube_acc_init(&A, C, nh_sum, P->K, JF, JF2)) { rc = -3; break; }
        if (P->threads > 1) {
            /* contiguous window ranges per thread; private bins merged in thread
             * order; per-cube sums are identical to the serial pass */
            const int T = P->threads;
            const long long WS = ORC_WAVE * A.K, nwin = (A.S + WS - 1) / WS;
            double** pw = (double**)calloc((size_t)T, sizeof(double*));
            uint32_t** pc = (uint32_t**)calloc((size_t)T, sizeof(uint32_t*));
            #pragma omp parallel num_threads(T)
            {
                int t = omp_get_thread_num();
                cube_acc Al = A;
                fill_ctx Fl = F;
                pw[t] = (double*)calloc((size_t)(d * n), sizeof(double));
                pc[t] = (uint32_t*)calloc((size_t)(d * n), sizeof(uint32_t));
                Fl.weights = pw[t];
                Fl.counts = pc[t];
                cube_sums_range(&Al, nwin * t / T, nwin * (t + 1) / T, sample_fn_impl, &Fl);
            }
            for (int t = 0; t < T; ++t) {
                for (int k = 0; k < d * n; ++k) { weights[k] += pw[t][k]; counts[k] += pc[t][k]; }
                free(pw[t]); free(pc[t]);
            }
            free(pw); free(pc);
            cube_fold(&A);
        } else {
            cube_sums_driver(&A, sample_fn_impl, &F);
        }
        cube_acc_free(&A);

        double sums[4];
        orc_results(C, V, P->beta, nh, JF, JF2, dh, sums);
        O->hist_I[it] = sums[0];
        O->hist_S2[it] = sums[1];
        O->hist_n[it] = S;
        if (it == 0) {
            if (O->JF_first) memcpy(O->JF_first, JF, sizeof(double) * (size_t)C);
            if (O->JF2_first) memcpy(O->JF2_first, JF2, sizeof(double) * (size_t)C);
            if (O->dh_first) memcpy(O->dh_first, dh, sizeof(double) * (size_t)C);
            if (O->w_first) memcpy(O->w_first, weights, sizeof(double) * (size_t)(d * n));
            if (O->c_first) memcpy(O->c_first, counts, sizeof(uint32_t) * (size_t)(d * n));
        }

        /* map refinement */
        for (int i = 0; i < d; ++i) {
            orc_refine_dim(n, P->alpha, &weights[i * n], &counts[i * n], &x_edges[i * E], x_new);
            memcpy(&x_edges[i * E], x_new, sizeof(double) * (size_t)E);
            for (int j = 0; j < n; ++j) dx[i * n + j] = x_edges[i * E + j + 1] - x_edges[i * E + j];
        }
        /* reallocation */
        orc_alloc(C, dh, sums[2], N, nh);
        if (it == 0) {
            if (O->nh_first) memcpy(O->nh_first, nh, sizeof(int) * (size_t)C);
            if (O->x_first) memcpy(O->x_first, x_edges, sizeof(double) * (size_t)(d * E));
        }
    }

    /* weighted average over iterations skip..iters-1 (cuda_wrapper.cu:750-766,869);
     * iterations of zero variance are exact (the library's rule, where the
     * reference's sums give inf / inf = NaN): their mean, error and chi2 0 */
    double zsum = 0.0;
    int nz = 0;
    for (int it = P->skip; it < P->iters; ++it)
        if (O->hist_S2[it] == 0.0) { zsum += O->hist_I[it]; ++nz; }
    if (nz > 0) {
        O->result = zsum / nz;
        O->error = 0.0;
        O->chi2_dof = 0.0;
    } else {
        double res = 0.0, sig = 0.0;
        int K = 0;
        for (int it = P->skip; it < P->iters; ++it) {
            res += O->hist_I[it] / O->hist_S2[it];
            sig += 1.0 / O->hist_S2[it];
            ++K;
        }
        if (K > 0) res /= sig;
        O->result = res;
        O->error = 1.0 / sqrt(sig);
        double chi2 = 0.0;
        for (int it = P->skip; it < P->iters; ++it) {
            double dd = O->hist_I[it] - res;
            chi2 += dd * dd / O->hist_S2[it];
        }
        O->chi2_dof = K > 1 ? chi2 / (K - 1) : 0.0;
    }
    if (O->x_edges) memcpy(O->x_edges, x_edges, sizeof(double) * (size_t)(d * E));
    if (O->nh) memcpy(O->nh, nh, sizeof(int) * (size_t)C);

    free(x_edges); free(x_new); free(dx); free(ef); free(dxf); free(weights); free(counts);
    free(nh); free(nh_sum); free(JF); free(JF2); free(dh);
    return rc;
}

/* Sample dump for a given map and allocation: cube, bin ids, x and jf of the
 * first dump_n samples of iteration `it` (bit-exact parity of the sampling
 * step).  x_edges: d*(n+1); nh: C. */
ORC_EXPORT int orc_sample_dump(const orc_params* P, const double* x_edges, const int* nh, int it,
                               long long dump_n, long long* cube, int* ids, double* x, double* jf) {
    const int d = P->d, n = P->nbins, E = n + 1;
    long long C;
    if (ipow_checked(P->nstrat, d, &C)) return -2;
    long long* nh_sum = (long long*)malloc(sizeof(long long) * (size_t)(C + 1));
    double* dx = (double*)malloc(sizeof(double) * (size_t)(d * n));
    float* ef = (float*)malloc(sizeof(float) * (size_t)(d * E));
    float* dxf = (float*)malloc(sizeof(float) * (size_t)(d * n));
    double* weights = (double*)calloc((size_t)(d * n), sizeof(double));
    uint32_t* counts = (uint32_t*)calloc((size_t)(d * n), sizeof(uint32_t));
    orc_prefix(C, nh, nh_sum);
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < n; ++j) dx[i * n + j] = x_edges[i * E + j + 1] - x_edges[i * E + j];
    build_f32_map(d, n, x_edges, ef, dxf);
    fill_ctx F;
    memset(&F, 0, sizeof F);
    F.P = P; F.C = C; F.it = it; F.x_edges = x_edges; F.dx = dx; F.ef = ef; F.dxf = dxf;
    F.weights = weights; F.counts = counts;
    for (int i = 0; i < P->n_iparams && i < 512; ++i) F.fparams[i] = (float)P->iparams[i];
    F.dump_n = dump_n; F.dump_ids = ids; F.dump_x = x; F.dump_jf = jf; F.dump_cube = cube;
    long long h = 0;
    for (long long s = 0; s < dump_n && s < nh_sum[C]; ++s) {
        while (nh_sum[h + 1] <= s) ++h;
        double a, b;
        sample_fn_impl(&F, s, h, &a, &b);
    }
    free(nh_sum); free(dx); free(ef); free(dxf); free(weights); free(counts);
    return 0;
}

/* Sizes of the structs, so the ctypes mirror can check its layout. */
ORC_EXPORT int orc_sizeof_params(void) { return (int)sizeof(orc_params); }
ORC_EXPORT int orc_sizeof_out(void) { return (int)sizeof(orc_out); }

/* Backward replay of one frozen iteration for the Gaussian mixture: the
 * samples of iteration `it` under map x_edges and allocation nh, with
 * parameters theta.  out[0] = I_k, out[1] = sigma_k^2, out[2..2+P) = dI_k/dtheta,
 * out[2+P..2+2P) = its stratified variance with the unbiased per-cube estimator
 * (same per-cube expressions as the GPU replay; cubes summed in index order). */
ORC_EXPORT int orc_grad_replay(const orc_params* P, const double* x_edges, const int* nh, int it,
                               const double* theta, double* out) {
    const int d = P->d, n = P->nbins, E = n + 1, NT = 8 + 4 * d;
    long long C;
    if (P->integrand != ORC_F_MIXTURE) return -1;
    if (ipow_checked(P->nstrat, d, &C)) return -2;
    const double V = pow(1.0 / P->nstrat, d);
    const uint32_t ns = (uint32_t)P->nstrat;
    double* dx = (double*)malloc(sizeof(double) * (size_t)(d * n));
    float* ef = (float*)malloc(sizeof(float) * (size_t)(d * E));
    float* dxf = (float*)malloc(sizeof(float) * (size_t)(d * n));
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < n; ++j) dx[i * n + j] = x_edges[i * E + j + 1] - x_edges[i * E + j];
    build_f32_map(d, n, x_edges, ef, dxf);
    float thf[512];
    for (int q = 0; q < NT; ++q) thf[q] = (float)theta[q];
    for (int q = 0; q < 2 + 2 * NT; ++q) out[q] = 0.0;
    long long s0 = 0;
    double g1[512], g2[512];
    for (long long h = 0; h < C; ++h) {
        const int cnt = nh[h];
        double sf = 0.0, sf2 = 0.0;
        for (int q = 0; q < NT; ++q) { g1[q] = 0.0; g2[q] = 0.0; }
        for (int e = 0; e < cnt; ++e) {
            const long long s = s0 + e;
            uint32_t words[68];
            for (int g = 0; g < (d + 3) / 4; ++g) {
                uint32_t ctr[4] = {(uint32_t)s, (uint32_t)((uint64_t)s >> 32), (uint32_t)it, (uint32_t)g};
                orc_philox(ctr, P->seed, P->stream, &words[4 * g]);
            }
            long long tmp = h;
            double jf, gr[512];
            if (P->precision == 0) {
                double x[64], jac = 1.0, dd[64];
                const double inv_ns = 1.0 / (double)ns;
                for (int i = 0; i < d; ++i) {
                    uint32_t k = (uint32_t)(tmp % ns); tmp /= ns;
                    uint32_t id, rem; uint64_t low;
                    orc_bin_of(words[i], k * (uint32_t)n, (uint32_t)n, ns, &id, &rem, &low);
                    double frac = ((double)rem + (double)low * 0x1p-33) * inv_ns;
                    x[i] = fma(dx[i * n + id], frac, x_edges[i * E + id]);
                    dd[i] = dx[i * n + id];
                }
                for (int i = 0; i < d; ++i) jac = jac * ((double)n * dd[i]);
                double f = mix64(x, d, theta, gr);
                jf = f * jac;
                for (int q = 0; q < NT; ++q) gr[q] = gr[q] * jac;
            } else {
                float x[64], jac = 1.0f, dd[64], grf[512];
                const float inv_ns = 1.0f / (float)ns;
                for (int i = 0; i < d; ++i) {
                    uint32_t k = (uint32_t)(tmp % ns); tmp /= ns;
                    uint32_t id, rem; uint64_t low;
                    orc_bin_of(words[i], k * (uint32_t)n, (uint32_t)n, ns, &id, &rem, &low);
                    float frac = ((float)rem + (float)(uint32_t)(low >> 9) * 0x1p-24f) * inv_ns;
                    x[i] = fmaf(dxf[i * n + id], frac, ef[i * E + id]);
                    dd[i] = dxf[i * n + id];
                }
                for (int i = 0; i < d; ++i) jac = jac * ((float)n * dd[i]);
                float f = mix32(x, d, thf, grf);
                jf = (double)(f * jac);
                for (int q = 0; q < NT; ++q) gr[q] = (double)(grf[q] * jac);
            }
            sf = sf + jf;
            sf2 = fma(jf, jf, sf2);
            for (int q = 0; q < NT; ++q) { g1[q] = g1[q] + gr[q]; g2[q] = fma(gr[q], gr[q], g2[q]); }
        }
        const double Ih = sf / cnt * V;
        const double Sig2 = (sf2 / cnt * V * V) - (Ih * Ih);
        out[0] += Ih;
        out[1] += Sig2 / cnt;
        const double vn = V / cnt, vvn = V * V / cnt, in1 = 1.0 / (cnt - 1);
        for (int q = 0; q < NT; ++q) {
            const double Gh = g1[q] * vn;
            const double Gv = g2[q] * vvn - Gh * Gh;
            out[2 + q] += Gh;
            out[2 + NT + q] += Gv * in1;
        }
        s0 += cnt;
    }
    free(dx); free(ef); free(dxf);
    return 0;
}
