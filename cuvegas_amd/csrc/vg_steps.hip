// Between-fill steps of the VEGAS+ iteration, compiled ahead of time for gfx950.
//
// Every step runs on the device so an iteration needs no host round trip
// (the reference synchronises the host 6+ times per iteration and launches
// 3d+15 kernels / CUB calls, cuda_wrapper.cu:575-773).  Reductions use one
// fixed order -- 1024-wide strided pairwise trees, repeated over partials --
// so the stratification weights D = sum d_h and the integer allocation n_h are
// reproducible bit for bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include "vg_math.h"
#include "vg_steps.h"

#define VG_TREE 1024
#define VG_FOLD_MAX 64  // cubes spanning more windows are folded by a tree (vg_bigfold)

// In-place strided pairwise tree over t[0..1023] (block of VG_TREE threads).
__device__ inline void vg_tree1024(double* t) {
    for (int stride = VG_TREE / 2; stride >= 1; stride >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < stride) t[threadIdx.x] = t[threadIdx.x] + t[threadIdx.x + stride];
    }
    __syncthreads();
}


// Exclusive scan of one int64 per thread over the block (wave64 shuffles, then
// the wave totals); *total receives the block sum.  wt: 16 int64 of LDS.
__device__ inline long long vg_block_scan_excl(long long v, long long* wt, long long* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    long long x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const long long y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    __syncthreads();
    if (lane == 63) wt[wid] = x;
    __syncthreads();
    if (wid == 0) {
        long long t = lane < nw ? wt[lane] : 0;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const long long y = __shfl_up(t, off, 64);
            if (lane >= off) t += y;
        }
        if (lane < nw) wt[lane] = t;
    }
    __syncthreads();
    const long long before = wid > 0 ? wt[wid - 1] : 0;
    *total = wt[nw - 1];
    return before + x - v;
}


// The 1024-wide strided pairwise tree (t[i] += t[i+512], then +256, ... +1)
// evaluated by 256 threads that each hold the elements t, t+256, t+512, t+768:
// the first two levels in registers, the next two through LDS, the last six
// with wave shuffles.  Bit-identical to the element-wise definition.  red:
// 256 doubles of LDS; the total is returned to thread 0.
__device__ inline double vg_tree1024_256(double v0, double v1, double v2, double v3, double* red) {
    const double a = v0 + v2;       // stride 512: i = t
    const double b = v1 + v3;       // stride 512: i = t + 256
    double x = a + b;               // stride 256: i = t
    __syncthreads();
    red[threadIdx.x] = x;
    __syncthreads();
    if (threadIdx.x < 128) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + 128];
    __syncthreads();
    if (threadIdx.x < 64) {
        x = red[threadIdx.x] + red[threadIdx.x + 64];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double y = __shfl_down(x, off, 64);
            x = x + y;  // lanes >= off compute values that are never read
        }
    }
    return x;
}

// Q trees of vg_tree1024_256 at once (value q of thread t: v[0..3][q] at
// elements t, t + 256, t + 512, t + 768), with three block barriers instead
// of 3 Q; each quantity's additions are those of vg_tree1024_256, so the
// totals are bit-identical.  red: Q * 256 doubles of LDS; tot[q] is valid in
// thread 0.
template <int Q>
__device__ inline void vg_tree1024_256q(const double (&v)[4][Q], double* red, double (&tot)[Q]) {
    double x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double a = v[0][q] + v[2][q];  // stride 512: i = t
        const double b = v[1][q] + v[3][q];  // stride 512: i = t + 256
        x[q] = a + b;                        // stride 256: i = t
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) red[q * 256 + threadIdx.x] = x[q];
    __syncthreads();
    if (threadIdx.x < 128) {
#pragma unroll
        for (int q = 0; q < Q; ++q) red[q * 256 + threadIdx.x] = red[q * 256 + threadIdx.x] + red[q * 256 + threadIdx.x + 128];
    }
    __syncthreads();
    if (threadIdx.x < 64) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double y = red[q * 256 + threadIdx.x] + red[q * 256 + threadIdx.x + 64];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const double z = __shfl_down(y, off, 64);
                y = y + z;  // lanes >= off compute values that are never read
            }
            tot[q] = y;
        }
    }
}

// True in every thread of the block that arrives last at *cnt among the
// gridDim.x blocks of this launch, with the other blocks' global writes
// visible to it (agent-scope release before the count, acquire after: on
// gfx950 each fence writes back / invalidates this XCD's L2, so this is for
// launches of a few dozen blocks).  The last block resets the counter, so
// it is zero again for the next launch without a memset.
__device__ inline bool vg_last_block(unsigned* cnt, int* flag) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned prev = atomicAdd(cnt, 1u);
        const int last = prev == gridDim.x - 1;
        if (last) atomicExch(cnt, 0u);
        *flag = last;
    }
    __syncthreads();
    const bool last = *flag != 0;
    if (last) __threadfence();
    return last;
}

// ---------------------------------------------------------------------------
// Per-cube results (reference `results`, cuda_wrapper.cu:26-49).  part[VG_NRES b + q]:
// block partial of q = 0: sum I_h, 1: sum sigma_h^2 / n_h, 2: sum d_h,
// 3: sum sigma_h^2 / (n_h - 1) (no reference counterpart).
// ---------------------------------------------------------------------------
// I_h, sigma_h^2 / n_h, d_h and sigma_h^2 / (n_h - 1) of a cube with sums
// j1 = sum J f, j2 = sum (J f)^2 over neval >= 2 samples, in the reference's
// expression order (cuda_wrapper.cu:33-47).  (A branch that wrote the
// divisions by n_h = 2 as products by 0.5 -- the same bits -- left the results
// launch unchanged, 125 against 123 us at 16.8 M cubes, and so did the same
// branch taken only by waves whose cubes all hold 2 samples (112-118 against
// 115-119 us after adaptation, where most waves take it): it is bound by
// memory and latency, not by the divide sequences.)
__device__ inline void vg_cube_terms(double j1, double j2, int neval, double V, double beta, double& sr,
                                     double& ss, double& dd, double& su) {
    const double Ih = j1 / neval * V;
    const double Sig2 = (j2 / neval * V * V) - (Ih * Ih);
    sr = Ih;
    ss = Sig2 / neval;
    su = Sig2 / (neval - 1);  // n_h >= 2 always
    const double u = V * j1 / neval;
    double t = V * V * j2 / neval - u * u;
    if (t < 0) t = 0;
    dd = vg_pow(t, beta);
}

// Results of the 1024 cubes of block b: d_h, and the block partials of
// (sum I_h, sum sigma_h^2 / n_h, sum d_h, sum sigma_h^2 / (n_h - 1)) in
// part[VG_NRES b + q].  256 threads, 4 cubes each; red: VG_NRES * 256 doubles
// of LDS.  The per-cube sums (complete: the window folds ran before) of all 4
// cubes are requested before any is evaluated, so a thread's loads overlap
// instead of waiting behind the previous cube's divides and pow.
__device__ inline void vg_results_block(long long b, int C, double V, double beta,
                                        const int* __restrict__ nh, const double2* __restrict__ JS,
                                        double* __restrict__ dh, double* __restrict__ part, double* red) {
    const long long base = b * VG_TREE + threadIdx.x;
    int ne[4];
    double2 js[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const long long h = base + 256 * j;
        ne[j] = 2;
        js[j] = make_double2(0.0, 0.0);
        if (h < C) { ne[j] = nh[h]; js[j] = JS[h]; }
    }
    double r[4][VG_NRES];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const long long h = base + 256 * j;
        double sr = 0.0, ss = 0.0, dd = 0.0, su = 0.0;
        if (h < C) {
            vg_cube_terms(js[j].x, js[j].y, ne[j], V, beta, sr, ss, dd, su);
            dh[h] = dd;
        }
        r[j][0] = sr; r[j][1] = ss; r[j][2] = dd; r[j][3] = su;
    }
    double tot[VG_NRES];
    vg_tree1024_256q<VG_NRES>(r, red, tot);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < VG_NRES; ++q) part[VG_NRES * b + q] = tot[q];
    }
}

__global__ void __launch_bounds__(256)
vg_results_kernel(int C, double V, double beta, const int* __restrict__ nh,
                  const double2* __restrict__ JS, double* __restrict__ dh, double* __restrict__ part) {
    __shared__ double red[VG_NRES * 256];
    vg_results_block(blockIdx.x, C, V, beta, nh, JS, dh, part, red);
}

// Cubes whose samples span more than VG_FOLD_MAX windows: the window partials
// win_last[wa], win_first[wa+1..wb] are summed with the 1024-wide tree, level
// after level, one block per cube (grid-stride over the list).
__device__ void vg_fold_tree(long long wa, long long m, const double2* __restrict__ win_first,
                             const double2* __restrict__ win_last, double* part, double* red,
                             double& o1, double& o2) {
    // level 1 from the window partials, later levels from part[] (2 x 2048 doubles:
    // m <= 2^21 windows, which holds for any allocation below 2^31 samples)
    long long cur = m;
    bool first = true;
    while (true) {
        const long long nch = (cur + VG_TREE - 1) / VG_TREE;
        for (long long c = 0; c < nch; ++c) {
            double v1[4], v2[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long long k = c * VG_TREE + threadIdx.x + 256 * j;
                if (k >= cur) { v1[j] = 0.0; v2[j] = 0.0; }
                else if (first) {
                    const double2 v = k == 0 ? win_last[wa] : win_first[wa + k];
                    v1[j] = v.x; v2[j] = v.y;
                } else { v1[j] = part[2 * k]; v2[j] = part[2 * k + 1]; }
            }
            __syncthreads();
            const double t1 = vg_tree1024_256(v1[0], v1[1], v1[2], v1[3], red);
            const double t2 = vg_tree1024_256(v2[0], v2[1], v2[2], v2[3], red);
            if (threadIdx.x == 0) { part[2 * c] = t1; part[2 * c + 1] = t2; }
            __syncthreads();
        }
        if (nch == 1) break;
        cur = nch;
        first = false;
    }
    o1 = part[0];
    o2 = part[1];
    __syncthreads();
}

// Blocks [0, VG_BIGFOLD_BLOCKS): the tree folds above.  The other blocks:
// cubes spanning 2..VG_FOLD_MAX windows, found at window starts -- the cube
// holding the first sample of window w (run_cube[64 w]) crosses from window
// w - 1 when it starts there; its partials are left-folded in window order,
// win_last[w - 1] + win_first[w] + ... + win_first[wb], into JS.
#define VG_BIGFOLD_BLOCKS 128
__global__ void __launch_bounds__(256)
vg_bigfold_kernel(const int* __restrict__ list, const int* __restrict__ count, int cap,
                  const long long* __restrict__ nh_sum, const int* __restrict__ run_cube,
                  const long long* __restrict__ s_total, const double2* __restrict__ win_first,
                  const double2* __restrict__ win_last, int ws_shift, double2* __restrict__ JS) {
    __shared__ double red[256];
    __shared__ double part[2 * 2048];
    if (blockIdx.x >= VG_BIGFOLD_BLOCKS) {
        const long long nwin = (*s_total + (1LL << ws_shift) - 1) >> ws_shift;
        const long long stride = (long long)(gridDim.x - VG_BIGFOLD_BLOCKS) * blockDim.x;
        for (long long w = 1 + (long long)(blockIdx.x - VG_BIGFOLD_BLOCKS) * blockDim.x + threadIdx.x;
             w < nwin; w += stride) {
            const int h = run_cube[w << 6];
            const long long st = nh_sum[h];
            if ((st >> ws_shift) != w - 1) continue;  // starts at window w, or folded from an earlier one
            const long long wb = (nh_sum[h + 1] - 1) >> ws_shift;
            if (wb - (w - 1) + 1 > VG_FOLD_MAX) continue;  // tree fold
            double2 v = win_last[w - 1];
            double j1 = v.x, j2 = v.y;
            for (long long u = w; u <= wb; ++u) {
                v = win_first[u];
                j1 = j1 + v.x;
                j2 = j2 + v.y;
            }
            JS[h] = make_double2(j1, j2);
        }
        return;
    }
    const int n = *count < cap ? *count : cap;
    for (int i = blockIdx.x; i < n; i += VG_BIGFOLD_BLOCKS) {
        const int h = list[i];
        const long long st = nh_sum[h], en = nh_sum[h + 1];
        const long long wa = st >> ws_shift, wb = (en - 1) >> ws_shift;
        double o1, o2;
        vg_fold_tree(wa, wb - wa + 1, win_first, win_last, part, red, o1, o2);
        if (threadIdx.x == 0) JS[h] = make_double2(o1, o2);
    }
}

hipError_t vg_launch_bigfold(hipStream_t st, const int* list, const int* count, int cap,
                             const long long* nh_sum, const int* run_cube, const long long* s_total,
                             const double2* win_first, const double2* win_last, int ws_shift,
                             double2* JS) {
    hipLaunchKernelGGL(vg_bigfold_kernel, dim3(VG_BIGFOLD_BLOCKS + 256), dim3(256), 0, st, list, count,
                       cap, nh_sum, run_cube, s_total, win_first, win_last, ws_shift, JS);
    return hipGetLastError();
}

// Remaining tree levels over the block partials, then the iteration record:
// hist[VG_HIST it + {0..4}] = {I_k, sigma_k^2, D, samples, unbiased sigma_k^2};
// *dsum = D.  Block g reduces chunk g (1024 block partials) into
// part2[VG_NRES g + q]; the last
// block to finish reduces part2 level by level, in place.  The sums are those
// of the sequential definition (chunks of 1024, strided pairwise trees,
// repeated over the partials).
__global__ void __launch_bounds__(256)
vg_finalize_kernel(long long nparts, const double* __restrict__ part, double* __restrict__ part2,
                   unsigned* __restrict__ sync, const long long* __restrict__ s_total,
                   const int* __restrict__ it_p, int hist_cap, double* __restrict__ hist,
                   double* __restrict__ dsum, int* __restrict__ err) {
    __shared__ double red[VG_NRES * 256];
    __shared__ int last;
    {
        double v[4][VG_NRES];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long i = (long long)blockIdx.x * VG_TREE + threadIdx.x + 256 * j;
#pragma unroll
            for (int q = 0; q < VG_NRES; ++q) v[j][q] = i < nparts ? part[VG_NRES * i + q] : 0.0;
        }
        double tot[VG_NRES];
        vg_tree1024_256q<VG_NRES>(v, red, tot);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < VG_NRES; ++q) part2[VG_NRES * blockIdx.x + q] = tot[q];
        }
    }
    if (!vg_last_block(sync, &last)) return;
    long long cur = gridDim.x;
    while (cur > 1) {
        const long long nch = (cur + VG_TREE - 1) / VG_TREE;
        for (long long c = 0; c < nch; ++c) {
            double v[4][VG_NRES];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long long i = c * VG_TREE + threadIdx.x + 256 * j;
#pragma unroll
                for (int q = 0; q < VG_NRES; ++q) v[j][q] = i < cur ? part2[VG_NRES * i + q] : 0.0;
            }
            __syncthreads();  // every read of this chunk precedes the writes below
            double tot[VG_NRES];
            vg_tree1024_256q<VG_NRES>(v, red, tot);
            if (threadIdx.x == 0) {
#pragma unroll
                for (int q = 0; q < VG_NRES; ++q) part2[VG_NRES * c + q] = tot[q];
            }
            __syncthreads();
        }
        cur = nch;
    }
    if (threadIdx.x == 0) {
        const int it = *it_p;
        *dsum = part2[2];
        if (it < hist_cap) {
            hist[VG_HIST * it + 0] = part2[0];
            hist[VG_HIST * it + 1] = part2[1];
            hist[VG_HIST * it + 2] = part2[2];
            hist[VG_HIST * it + 3] = (double)*s_total;
            hist[VG_HIST * it + 4] = part2[3];
        } else {
            atomicOr(err, VG_ERR_HIST);
        }
    }
}

// ---------------------------------------------------------------------------
// Reallocation (reference `updateNh`, cuda_wrapper.cu:118-132):
//   n_h = max(2, (int)(d_h / D * N))        (mode 1)
//   n_h = nh_const                          (mode 0, initial allocation)
// plus the per-block sums feeding the prefix sum.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
vg_alloc_kernel(int C, int mode, int nh_const, const double* __restrict__ dh,
                const double* __restrict__ dsum, long long N, int* __restrict__ nh,
                long long* __restrict__ bsum) {
    __shared__ long long red[4];
    const long long base = (long long)blockIdx.x * VG_TREE + threadIdx.x;
    const double D = mode ? *dsum : 1.0;
    long long v = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const long long h = base + 256 * j;
        if (h < C) {
            int n;
            if (mode == 0) {
                n = nh_const;
            } else {
                const double f = dh[h] / D;
                n = vg_sat_int(f * (double)N);
                n = n < 2 ? 2 : n;
            }
            nh[h] = n;
            v += n;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Exclusive scan of the block sums.  Block g scans its 1024 block sums into
// boff (offsets within group g) and writes the group total gsum[g]; the last
// block to finish scans the group totals into goff.  The cube-level scan adds
// boff[b] + goff[b / 1024].  Total -> *s_total, capped at s_cap (error flag
// set if the allocation exceeds the sample buffers).  Optionally advances the
// iteration counter (end of an iteration).
__global__ void __launch_bounds__(VG_TREE)
vg_scan_blocks_kernel(long long nb, const long long* __restrict__ bsum, long long* __restrict__ boff,
                      long long* __restrict__ gsum, long long* __restrict__ goff,
                      unsigned* __restrict__ sync, long long* __restrict__ s_total, long long s_cap,
                      int* __restrict__ it_p, int advance, int* __restrict__ err,
                      int* __restrict__ counts) {
    __shared__ long long wt[VG_TREE / 64];
    __shared__ int last;
    {
        const long long b = (long long)blockIdx.x * VG_TREE + threadIdx.x;
        long long tot;
        const long long ex = vg_block_scan_excl(b < nb ? bsum[b] : 0, wt, &tot);
        if (b < nb) boff[b] = ex;
        if (threadIdx.x == 0) gsum[blockIdx.x] = tot;
    }
    if (!vg_last_block(sync, &last)) return;
    const long long ng = gridDim.x;
    long long carry = 0;
    for (long long g0 = 0; g0 < ng; g0 += VG_TREE) {
        const long long g = g0 + threadIdx.x;
        long long tot;
        const long long ex = vg_block_scan_excl(g < ng ? gsum[g] : 0, wt, &tot);
        if (g < ng) goff[g] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        long long total = carry;
        if (total > s_cap) { atomicOr(err, VG_ERR_CAPACITY); total = s_cap; }
        *s_total = total;
        if (advance) *it_p = *it_p + 1;
        counts[0] = 0;
        counts[1] = 0;
    }
}

// Per-block local scan of n_h -> nh_sum (exclusive, nh_sum[C] = total, every
// entry capped at the sample total), the run table -- run_cube[r] = the cube
// holding sample r * K (K = 2^k_shift) -- and the list of cubes whose samples
// span more than VG_FOLD_MAX sampling windows.  256 threads own 4 consecutive
// cubes each (16-byte load, 2 x 16-byte stores).  A cube writes its own runs
// when it holds at most VG_OWN_RUNS of them; larger cubes go to big_list in
// chunks of at most VG_BIG_CHUNK runs (r0, n << 32 | h), which
// vg_big_runs_kernel spreads over the whole grid, so a cube holding millions
// of samples does not hold up one block.  counts[0]: fold list length,
// counts[1]: big list length (both cleared by vg_scan_blocks_kernel).
#define VG_OWN_RUNS 16
#define VG_BIG_CHUNK 8192
__global__ void __launch_bounds__(256)
vg_scan_cubes_kernel(int C, const int* __restrict__ nh, const long long* __restrict__ boff,
                     const long long* __restrict__ goff, const long long* __restrict__ s_total, int k_shift,
                     long long run_cap, long long* __restrict__ nh_sum, int* __restrict__ run_cube,
                     int* __restrict__ fold_list, int* __restrict__ counts, int fold_cap,
                     long long* __restrict__ big_list, int big_cap) {
    __shared__ long long wt[4];
    const long long base = (long long)blockIdx.x * VG_TREE;
    const long long h0 = base + 4 * threadIdx.x;
    const bool full = h0 + 4 <= C;
    long long v[4];
    long long own = 0;
    if (full) {
        const int4 q = *reinterpret_cast<const int4*>(nh + h0);  // h0 % 4 == 0: aligned
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = h0 + j < C ? nh[h0 + j] : 0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) own += v[j];
    long long total;
    const long long S = *s_total;
    const long long ex = vg_block_scan_excl(own, wt, &total) + boff[blockIdx.x] + goff[blockIdx.x / VG_TREE];
    long long e[5];
    e[0] = ex;
#pragma unroll
    for (int j = 0; j < 4; ++j) e[j + 1] = e[j] + v[j];
#pragma unroll
    for (int j = 0; j < 5; ++j) e[j] = e[j] < S ? e[j] : S;
    if (full) {
        longlong2* o = reinterpret_cast<longlong2*>(nh_sum + h0);
        o[0] = make_longlong2(e[0], e[1]);
        o[1] = make_longlong2(e[2], e[3]);
    }
    const long long kmask = (1LL << k_shift) - 1;
    long long r0[4], nr[4];
    long long nchunk = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        r0[j] = (e[j] + kmask) >> k_shift;
        long long r1 = (e[j + 1] + kmask) >> k_shift;
        if (r1 > run_cap) r1 = run_cap;
        nr[j] = h0 + j < C ? r1 - r0[j] : 0;
        if (nr[j] > VG_OWN_RUNS) nchunk += (nr[j] + VG_BIG_CHUNK - 1) / VG_BIG_CHUNK;
    }
    // list slots: one global atomic per block that has big cubes
    __shared__ long long slot0;
    long long nblock;
    long long slot = vg_block_scan_excl(nchunk, wt, &nblock);
    if (nblock > 0) {
        if (threadIdx.x == 0) slot0 = atomicAdd(counts + 1, (int)nblock);
        __syncthreads();
        slot += slot0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const long long h = h0 + j;
        if (h < C) {
            if (!full) nh_sum[h] = e[j];
            if (h == C - 1) nh_sum[C] = S;
            const long long n = nr[j];
            if (n > VG_OWN_RUNS) {
                for (long long c = 0; c * VG_BIG_CHUNK < n; ++c, ++slot) {
                    const long long m = n - c * VG_BIG_CHUNK < VG_BIG_CHUNK ? n - c * VG_BIG_CHUNK : VG_BIG_CHUNK;
                    const long long rc = r0[j] + c * VG_BIG_CHUNK;
                    if (slot < big_cap) {
                        big_list[2 * slot] = rc;
                        big_list[2 * slot + 1] = (m << 32) | h;
                    } else {  // never with big_cap = vg_big_list_cap(run_cap); kept correct anyway
                        for (long long r = rc; r < rc + m; ++r) run_cube[r] = (int)h;
                    }
                }
            } else {
                for (long long r = r0[j]; r < r0[j] + n; ++r) run_cube[r] = (int)h;
            }
            const long long st = e[j], en = e[j + 1];
            if (en > st && ((en - 1) >> (k_shift + 6)) - (st >> (k_shift + 6)) + 1 > VG_FOLD_MAX) {
                const int fs = atomicAdd(counts, 1);
                if (fs < fold_cap) fold_list[fs] = (int)h;
            }
        }
    }
}

// Runs of the cubes in big_list: one wave per chunk (grid-stride over waves).
__global__ void __launch_bounds__(256)
vg_big_runs_kernel(const long long* __restrict__ big_list, const int* __restrict__ counts, int big_cap,
                   int* __restrict__ run_cube) {
    const int n = counts[1] < big_cap ? counts[1] : big_cap;
    const int lane = threadIdx.x & 63;
    const int nw = gridDim.x * 4;
    for (int e = blockIdx.x * 4 + (threadIdx.x >> 6); e < n; e += nw) {
        const long long r0 = big_list[2LL * e];
        const long long w = big_list[2LL * e + 1];
        const int h = (int)(w & 0xffffffffLL);
        const int m = (int)(w >> 32);
        for (int i = lane; i < m; i += 64) run_cube[r0 + i] = h;
    }
}

long long vg_big_list_cap(long long run_cap) {
    // every listed cube holds > VG_OWN_RUNS runs, and chunking adds at most
    // one entry per VG_BIG_CHUNK runs
    return run_cap / (VG_OWN_RUNS + 1) + run_cap / VG_BIG_CHUNK + 2;
}

// ---------------------------------------------------------------------------
// Map: uniform initialisation (cuda_wrapper.cu:464-473) and refinement
// (normalizeWeights / smoothWeights / setMap / updateXEdges / updateDxEdges,
// cuda_wrapper.cu:51-90,134-191 and the loop at :719-746).  One block per
// dimension.  Also rebuilds the packed sampling tables and clears the bin
// accumulators for the next fill.
// ---------------------------------------------------------------------------

__device__ void vg_write_tables(int n, const double* __restrict__ x, double2* __restrict__ map64,
                                float2* __restrict__ map32) {
    // (edge, n * width) when n is a power of two (an exact scaling: the
    // sampling kernels read the Jacobian factor n * width directly and scale
    // the in-bin offset by 1/n instead), else (edge, width)
    const double ws = (n & (n - 1)) == 0 ? (double)n : 1.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const double a = x[k], b = x[k + 1];
        map64[k] = make_double2(a, (b - a) * ws);
        const float fa = (float)a, fb = (float)b;
        map32[k] = make_float2(fa, (fb - fa) * (float)ws);
    }
}

__global__ void __launch_bounds__(VG_TREE)
vg_init_map_kernel(int n, double* __restrict__ x_edges, double2* __restrict__ map64,
                   float2* __restrict__ map32, double* __restrict__ weights,
                   unsigned* __restrict__ counts) {
    const int dim = blockIdx.x;
    double* x = x_edges + (long long)dim * (n + 1);
    const double step = 1.0 / n;
    for (int j = threadIdx.x; j <= n; j += blockDim.x) x[j] = j * step;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        weights[(long long)dim * n + k] = 0.0;
        counts[(long long)dim * n + k] = 0u;
    }
    __syncthreads();
    vg_write_tables(n, x, map64 + (long long)dim * n, map32 + (long long)dim * n);
}

// Strided-pairwise tree sum of a[0..m), result broadcast; the same order as
// the per-cube reductions.  part: scratch of (m+1023)/1024 entries.  NT = 1024
// threads (t: 1024 doubles of LDS) or 256 threads holding 4 elements each
// (t: 256 doubles; vg_tree1024_256, bit-identical).
template <int NT>
__device__ double vg_block_tree_sum(const double* __restrict__ a, int m, double* __restrict__ part,
                                    double* t) {
    int cur = m;
    const double* src = a;
    while (true) {
        const int nch = (cur + VG_TREE - 1) / VG_TREE;
        for (int c = 0; c < nch; ++c) {
            if (NT == VG_TREE) {
                const int j = c * VG_TREE + threadIdx.x;
                __syncthreads();
                t[threadIdx.x] = j < cur ? src[j] : 0.0;
                vg_tree1024(t);
                if (threadIdx.x == 0) part[c] = t[0];
            } else {
                double v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int j = c * VG_TREE + threadIdx.x + 256 * q;
                    v[q] = j < cur ? src[j] : 0.0;
                }
                const double tot = vg_tree1024_256(v[0], v[1], v[2], v[3], t);
                if (threadIdx.x == 0) part[c] = tot;
            }
        }
        __syncthreads();
        if (nch == 1) break;
        cur = nch;
        src = part;
    }
    __syncthreads();
    const double r = part[0];
    __syncthreads();
    return r;
}

// Working arrays of one dimension's refinement.
struct vg_refine_ws {
    double* w;     // n   normalised bin weights
    double* d;     // n   smoothed, damped weights
    double* cum;   // n   their running sum
    double* xo;    // n+1 edges before the update
    double* part;  // n/1024 + 2 tree partials
};

// The arrays laid out consecutively from base (LDS or global).
__device__ inline vg_refine_ws vg_refine_ws_at(double* base, int n) {
    vg_refine_ws ws;
    ws.w = base;
    ws.d = ws.w + n;
    ws.cum = ws.d + n;
    ws.xo = ws.cum + n;
    ws.part = ws.xo + n + 1;
    return ws;
}

// Refinement of one dimension: normalise the bin sums, smooth, damp, cumulate,
// invert (reference normalizeWeights / smoothWeights / setMap / updateXEdges /
// updateDxEdges, cuda_wrapper.cu:51-90,134-191), with a block of NT threads.
template <int NT>
__device__ void vg_refine_dim(int n, double alpha, double* __restrict__ x, double* __restrict__ wb,
                              unsigned* __restrict__ cb, vg_refine_ws ws, double2* __restrict__ map64,
                              float2* __restrict__ map32, double* t) {
    const int E = n + 1;
    double* __restrict__ w = ws.w;
    double* __restrict__ d = ws.d;
    double* __restrict__ cum = ws.cum;
    double* __restrict__ xo = ws.xo;
    double* __restrict__ part = ws.part;

    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const unsigned c = cb[k];
        w[k] = c != 0u ? wb[k] / (double)c : wb[k];
        wb[k] = 0.0;
        cb[k] = 0u;
    }
    for (int j = threadIdx.x; j < E; j += blockDim.x) xo[j] = x[j];
    __syncthreads();
    const double dsum = vg_block_tree_sum<NT>(w, n, part, t);
    if (dsum > 0.0) {
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            double v;
            if (k == 0) v = (7.0 * w[k] + w[k + 1]) / (8.0 * dsum);
            else if (k == n - 1) v = (w[k - 1] + 7.0 * w[k]) / (8.0 * dsum);
            else v = (w[k - 1] + 6.0 * w[k] + w[k + 1]) / (8.0 * dsum);
            if (v != 0) v = vg_pow((v - 1.0) / vg_log(v), alpha);
            d[k] = v;
        }
        __syncthreads();
        const double S = vg_block_tree_sum<NT>(d, n, part, t);
        if (threadIdx.x == 0) {  // sequential inclusive sum, loads batched ahead of the adds
            double acc = 0.0;
            int k = 0;
            for (; k + 8 <= n; k += 8) {
                double v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = d[k + j];
#pragma unroll
                for (int j = 0; j < 8; ++j) { acc = acc + v[j]; cum[k + j] = acc; }
            }
            for (; k < n; ++k) { acc = acc + d[k]; cum[k] = acc; }
        }
        __syncthreads();
        const double delta = S / n;
        for (int j = threadIdx.x; j < n - 1; j += blockDim.x) {
            const int target = j + 1;
            int lo = 0, hi = n;  // first k with (int)(cum_k / delta) >= target
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (vg_sat_int(cum[mid] / delta) >= target) hi = mid; else lo = mid + 1;
            }
            const int k = lo < n ? lo : n - 1;
            const double a = target * delta - (k > 0 ? cum[k - 1] : 0.0);
            x[j + 1] = xo[k] + a / d[k] * (xo[k + 1] - xo[k]);
        }
    }
    __syncthreads();
    vg_write_tables(n, x, map64, map32);
}

// One dimension per block; working arrays in dynamic LDS when they fit
// (use_lds != 0), else in the global scratch.
__global__ void __launch_bounds__(VG_TREE)
vg_refine_kernel(int n, double alpha, int use_lds, double* __restrict__ x_edges,
                 double* __restrict__ weights, unsigned* __restrict__ counts,
                 double* __restrict__ scratch, double2* __restrict__ map64,
                 float2* __restrict__ map32) {
    __shared__ double t[VG_TREE];
    extern __shared__ double dyn[];
    const int dim = blockIdx.x;
    double* base = use_lds ? dyn : scratch + (long long)dim * (4LL * n + 1 + n / VG_TREE + 2);
    vg_refine_dim<VG_TREE>(n, alpha, x_edges + (long long)dim * (n + 1), weights + (long long)dim * n,
                           counts + (long long)dim * n, vg_refine_ws_at(base, n), map64 + (long long)dim * n,
                           map32 + (long long)dim * n, t);
}

// Per-cube results in blocks rd.. and, beside them, the map refinement of
// dimension b in block b < rd (the two read and write disjoint buffers, and
// the refinement of a few dimensions is latency-bound, so it runs in the
// shadow of the results).  The refinement keeps d, cum and the tree partials
// in dynamic LDS (2n + n/1024 + 2 doubles) and w, xo in the global scratch.
__global__ void __launch_bounds__(256)
vg_results_refine_kernel(int C, double V, double beta, const int* __restrict__ nh,
                         const double2* __restrict__ JS, double* __restrict__ dh, double* __restrict__ part,
                         int rd, int n, double alpha, double* __restrict__ x_edges,
                         double* __restrict__ weights, unsigned* __restrict__ counts,
                         double* __restrict__ scratch, double2* __restrict__ map64,
                         float2* __restrict__ map32) {
    __shared__ double red[256];
    extern __shared__ double dyn[];
    if ((int)blockIdx.x < rd) {
        const int dim = blockIdx.x;
        vg_refine_ws ws = vg_refine_ws_at(scratch + (long long)dim * (4LL * n + 1 + n / VG_TREE + 2), n);
        ws.d = dyn;
        ws.cum = dyn + n;
        ws.part = dyn + 2 * n;
        vg_refine_dim<256>(n, alpha, x_edges + (long long)dim * (n + 1), weights + (long long)dim * n,
                           counts + (long long)dim * n, ws, map64 + (long long)dim * n,
                           map32 + (long long)dim * n, red);
        return;
    }
    // the per-cube results blocks take their tree buffer from the dynamic LDS
    // (sized for both uses)
    vg_results_block(blockIdx.x - rd, C, V, beta, nh, JS, dh, part, dyn);
}

long long vg_results_refine_lds(int n) {
    const long long refine = 8LL * (2LL * n + n / VG_TREE + 2), tree = 8LL * VG_NRES * 256;
    return refine > tree ? refine : tree;
}

long long vg_refine_scratch_doubles(int d, int n) {
    return (long long)d * (4LL * n + 1 + n / VG_TREE + 2);
}

// ---------------------------------------------------------------------------
// Host-side launch helpers (all on one stream, no synchronisation).
// ---------------------------------------------------------------------------

hipError_t vg_launch_results(hipStream_t st, int C, double V, double beta, const int* nh,
                             const double2* JS, double* dh, double* part) {
    const unsigned nb = (unsigned)((C + VG_TREE - 1) / VG_TREE);
    hipLaunchKernelGGL(vg_results_kernel, dim3(nb), dim3(256), 0, st, C, V, beta, nh, JS, dh, part);
    return hipGetLastError();
}

hipError_t vg_launch_results_refine(hipStream_t st, int C, double V, double beta, const int* nh,
                                    const double2* JS, double* dh, double* part, int d, int n, double alpha,
                                    double* x_edges, double* weights, unsigned* counts, double* scratch,
                                    double2* map64, float2* map32) {
    const unsigned nb = (unsigned)((C + VG_TREE - 1) / VG_TREE);
    hipLaunchKernelGGL(vg_results_refine_kernel, dim3(nb + d), dim3(256), (size_t)vg_results_refine_lds(n), st,
                       C, V, beta, nh, JS, dh, part, d, n, alpha, x_edges, weights, counts, scratch, map64,
                       map32);
    return hipGetLastError();
}

hipError_t vg_launch_finalize(hipStream_t st, long long nparts, double* part, unsigned* sync,
                              const long long* s_total, const int* it_p, int hist_cap, double* hist,
                              double* dsum, int* err) {
    const long long ng = (nparts + VG_TREE - 1) / VG_TREE;
    hipLaunchKernelGGL(vg_finalize_kernel, dim3((unsigned)ng), dim3(256), 0, st, nparts, part,
                       part + VG_NRES * nparts, sync, s_total, it_p, hist_cap, hist, dsum, err);
    return hipGetLastError();
}

hipError_t vg_launch_alloc(hipStream_t st, int C, int mode, int nh_const, const double* dh,
                           const double* dsum, long long N, int* nh, long long* bsum) {
    const unsigned nb = (unsigned)((C + VG_TREE - 1) / VG_TREE);
    hipLaunchKernelGGL(vg_alloc_kernel, dim3(nb), dim3(256), 0, st, C, mode, nh_const, dh, dsum,
                       N, nh, bsum);
    return hipGetLastError();
}

hipError_t vg_launch_scan(hipStream_t st, int C, const int* nh, long long* bsum, long long* boff,
                          unsigned* sync, long long* s_total, long long s_cap, int* it_p, int advance,
                          int k_shift, long long run_cap, long long* nh_sum, int* run_cube, int* err,
                          int* fold_list, int* counts, int fold_cap, long long* big_list, int big_cap) {
    const long long nb = (C + VG_TREE - 1) / VG_TREE;
    const long long ng = (nb + VG_TREE - 1) / VG_TREE;
    hipLaunchKernelGGL(vg_scan_blocks_kernel, dim3((unsigned)ng), dim3(VG_TREE), 0, st, nb, bsum, boff,
                       bsum + nb, boff + nb, sync, s_total, s_cap, it_p, advance, err, counts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (!run_cube) run_cap = 0;
    hipLaunchKernelGGL(vg_scan_cubes_kernel, dim3((unsigned)nb), dim3(256), 0, st, C, nh, boff, boff + nb,
                       s_total, k_shift, run_cap, nh_sum, run_cube, fold_list, counts, fold_cap, big_list,
                       big_cap);
    e = hipGetLastError();
    if (e != hipSuccess || run_cap <= 0) return e;
    hipLaunchKernelGGL(vg_big_runs_kernel, dim3(2048), dim3(256), 0, st, big_list, counts, big_cap, run_cube);
    return hipGetLastError();
}

hipError_t vg_launch_init_map(hipStream_t st, int d, int n, double* x_edges, double2* map64,
                              float2* map32, double* weights, unsigned* counts) {
    hipLaunchKernelGGL(vg_init_map_kernel, dim3(d), dim3(VG_TREE), 0, st, n, x_edges, map64, map32,
                       weights, counts);
    return hipGetLastError();
}

hipError_t vg_launch_refine(hipStream_t st, int d, int n, double alpha, double* x_edges,
                            double* weights, unsigned* counts, double* scratch, double2* map64,
                            float2* map32) {
    const size_t lds = sizeof(double) * (size_t)(4LL * n + 1 + n / VG_TREE + 2);
    const int use_lds = lds <= 128 * 1024;
    if (use_lds && lds > 48 * 1024) {
        static bool raised = false;
        if (!raised) {
            hipError_t e = hipFuncSetAttribute((const void*)vg_refine_kernel,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024);
            if (e != hipSuccess) return e;
            raised = true;
        }
    }
    hipLaunchKernelGGL(vg_refine_kernel, dim3(d), dim3(VG_TREE), use_lds ? lds : 0, st, n, alpha,
                       use_lds, x_edges, weights, counts, scratch, map64, map32);
    return hipGetLastError();
}

__global__ void vg_write_tables_kernel(int n, const double* __restrict__ x_edges,
                                       double2* __restrict__ map64, float2* __restrict__ map32) {
    const int dim = blockIdx.x;
    vg_write_tables(n, x_edges + (long long)dim * (n + 1), map64 + (long long)dim * n,
                    map32 + (long long)dim * n);
}

hipError_t vg_launch_tables(hipStream_t st, int d, int n, const double* x_edges, double2* map64,
                            float2* map32) {
    hipLaunchKernelGGL(vg_write_tables_kernel, dim3(d), dim3(256), 0, st, n, x_edges, map64, map32);
    return hipGetLastError();
}

// Block sums of an existing n_h (no reallocation), then the prefix layout.
__global__ void __launch_bounds__(VG_TREE)
vg_bsum_kernel(int C, const int* __restrict__ nh, long long* __restrict__ bsum) {
    __shared__ long long t[VG_TREE];
    const long long h = (long long)blockIdx.x * VG_TREE + threadIdx.x;
    t[threadIdx.x] = h < C ? nh[h] : 0;
    for (int stride = VG_TREE / 2; stride >= 1; stride >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < stride) t[threadIdx.x] += t[threadIdx.x + stride];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = t[0];
}

hipError_t vg_launch_layout(hipStream_t st, int C, const int* nh, long long* bsum, long long* boff,
                            unsigned* sync, long long* s_total, long long s_cap, int* it_p, int k_shift,
                            long long run_cap, long long* nh_sum, int* run_cube, int* err, int* fold_list,
                            int* counts, int fold_cap, long long* big_list, int big_cap) {
    const unsigned nb = (unsigned)((C + VG_TREE - 1) / VG_TREE);
    hipLaunchKernelGGL(vg_bsum_kernel, dim3(nb), dim3(VG_TREE), 0, st, C, nh, bsum);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return vg_launch_scan(st, C, nh, bsum, boff, sync, s_total, s_cap, it_p, 0, k_shift, run_cap, nh_sum,
                          run_cube, err, fold_list, counts, fold_cap, big_list, big_cap);
}

// Copy the map edges and allocation that the coming sampling pass will use
// into slot (it - skip) of the record, for iterations skip <= it < cap.
__global__ void vg_snapshot_kernel(const int* __restrict__ it_p, int skip, int cap, int C, int nx,
                                   const double* __restrict__ x_edges, const int* __restrict__ nh,
                                   double* __restrict__ snap_x, int* __restrict__ snap_nh) {
    const int it = *it_p;
    if (it < skip || it >= cap) return;
    const long long slot = it - skip;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < C;
         i += (long long)gridDim.x * blockDim.x)
        snap_nh[slot * C + i] = nh[i];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nx;
         i += (long long)gridDim.x * blockDim.x)
        snap_x[slot * nx + i] = x_edges[i];
}

hipError_t vg_launch_snapshot(hipStream_t st, const int* it_p, int skip, int cap, int C, int nx,
                              const double* x_edges, const int* nh, double* snap_x, int* snap_nh) {
    const long long work = C > nx ? C : nx;
    unsigned grid = (unsigned)((work + 255) / 256);
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(vg_snapshot_kernel, dim3(grid), dim3(256), 0, st, it_p, skip, cap, C, nx,
                       x_edges, nh, snap_x, snap_nh);
    return hipGetLastError();
}

// ===========================================================================
// Batched steps: many independent integrals (items) per launch.  Each item
// keeps its own cube range, block range, map and bins; every per-item
// reduction uses the same 1024-wide trees as the single-integral kernels, so
// an item reproduces a single-integral run bit for bit.
// ===========================================================================

// Per-cube results of every item (blocks n_refine..), from complete per-cube
// sums (vg_bbigfold_kernel folded the window partials before), and beside
// them the map refinement of (item, dimension) rd_item[b], rd_dim[b] in block
// b < n_refine, 256 threads, working arrays in dynamic LDS
// (4n + 1 + n/1024 + 2 doubles).
__global__ void __launch_bounds__(256)
vg_bresults_kernel(const vg_bitem* __restrict__ items, const int* __restrict__ blk_item,
                   const int* __restrict__ nh, const double2* __restrict__ JS, double beta,
                   double* __restrict__ dh,
                   double* __restrict__ part, int n_refine, const int* __restrict__ rd_item,
                   const int* __restrict__ rd_dim, int n, double alpha, double* __restrict__ x_all,
                   double* __restrict__ weights, unsigned* __restrict__ counts,
                   double2* __restrict__ map64, float2* __restrict__ map32) {
    __shared__ double red[256];
    extern __shared__ double dyn[];
    if ((int)blockIdx.x < n_refine) {
        const vg_bitem R = items[rd_item[blockIdx.x]];
        const int dim = rd_dim[blockIdx.x];
        const long long mo = R.map_off + (long long)dim * n;
        vg_refine_dim<256>(n, alpha, x_all + R.xe_off + (long long)dim * (n + 1), weights + mo, counts + mo,
                           vg_refine_ws_at(dyn, n), map64 + mo, map32 + mo, red);
        return;
    }
    const long long bb = blockIdx.x - n_refine;
    const vg_bitem I = items[blk_item[bb]];
    const long long base = (bb - I.blk_off) * VG_TREE + threadIdx.x;
    // the sums of all 4 cubes are requested before any is evaluated
    int ne[4];
    double j1[4], j2[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const long long h = base + 256 * j;
        ne[j] = 2; j1[j] = 0.0; j2[j] = 0.0;
        if (h < I.C) {
            ne[j] = nh[I.cube_off + h];
            const double2 v = JS[I.cube_off + h];
            j1[j] = v.x;
            j2[j] = v.y;
        }
    }
    double r[4][VG_NRES];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const long long h = base + 256 * j;
        double sr = 0.0, ss = 0.0, dd = 0.0, su = 0.0;
        if (h < I.C) {
            vg_cube_terms(j1[j], j2[j], ne[j], I.V, beta, sr, ss, dd, su);
            dh[I.cube_off + h] = dd;
        }
        r[j][0] = sr; r[j][1] = ss; r[j][2] = dd; r[j][3] = su;
    }
    double tot[VG_NRES];
    vg_tree1024_256q<VG_NRES>(r, dyn, tot);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < VG_NRES; ++q) part[VG_NRES * bb + q] = tot[q];
    }
}

// One block per item: remaining tree levels over the item's block partials
// (in place), then hist[(item * cap + it) * VG_HIST + q] and dsum[item].
__global__ void __launch_bounds__(256)
vg_bfinalize_kernel(const vg_bitem* __restrict__ items, double* __restrict__ part_all,
                    const long long* __restrict__ s_item, const int* __restrict__ it_p, int hist_cap,
                    double* __restrict__ hist, double* __restrict__ dsum, int* __restrict__ err) {
    __shared__ double red[VG_NRES * 256];
    const int item = blockIdx.x;
    const vg_bitem I = items[item];
    double* part = part_all + VG_NRES * I.blk_off;
    long long cur = I.nblk;
    while (cur > 1) {
        const long long nch = (cur + VG_TREE - 1) / VG_TREE;
        for (long long c = 0; c < nch; ++c) {
            double v[4][VG_NRES];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long long i = c * VG_TREE + threadIdx.x + 256 * j;
#pragma unroll
                for (int q = 0; q < VG_NRES; ++q) v[j][q] = i < cur ? part[VG_NRES * i + q] : 0.0;
            }
            __syncthreads();
            double tot[VG_NRES];
            vg_tree1024_256q<VG_NRES>(v, red, tot);
            if (threadIdx.x == 0) {
#pragma unroll
                for (int q = 0; q < VG_NRES; ++q) part[VG_NRES * c + q] = tot[q];
            }
            __syncthreads();
        }
        cur = nch;
    }
    if (threadIdx.x == 0) {
        const int it = *it_p;
        dsum[item] = part[2];
        if (it < hist_cap) {
            double* hrow = hist + ((long long)item * hist_cap + it) * VG_HIST;
            hrow[0] = part[0];
            hrow[1] = part[1];
            hrow[2] = part[2];
            hrow[3] = (double)s_item[item];
            hrow[4] = part[3];
        } else {
            atomicOr(err, VG_ERR_HIST);
        }
    }
}

__global__ void __launch_bounds__(256)
vg_balloc_kernel(const vg_bitem* __restrict__ items, const int* __restrict__ blk_item, int mode,
                 const double* __restrict__ dh, const double* __restrict__ dsum,
                 int* __restrict__ nh, long long* __restrict__ bsum) {
    __shared__ long long red[4];
    const int item = blk_item[blockIdx.x];
    const vg_bitem I = items[item];
    const long long lb = blockIdx.x - I.blk_off;
    const long long base = lb * VG_TREE + threadIdx.x;
    const double D = mode ? dsum[item] : 1.0;
    long long v = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const long long h = base + 256 * j;
        if (h < I.C) {
            int n;
            if (mode == 0) {
                n = I.nh0;
            } else {
                const double f = dh[I.cube_off + h] / D;
                n = vg_sat_int(f * (double)I.N);
                n = n < 2 ? 2 : n;
            }
            nh[I.cube_off + h] = n;
            v += n;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// One block per item: exclusive scan of the item's block sums; total ->
// s_item[item] (capped); item 0 advances the iteration counter.
__global__ void __launch_bounds__(VG_TREE)
vg_bscan_blocks_kernel(const vg_bitem* __restrict__ items, const long long* __restrict__ bsum,
                       long long* __restrict__ boff, long long* __restrict__ s_item,
                       int* __restrict__ it_p, int advance, int* __restrict__ err,
                       int* __restrict__ counts) {
    __shared__ long long wt[VG_TREE / 64];
    const int item = blockIdx.x;
    const vg_bitem I = items[item];
    const long long nb = I.nblk, off = I.blk_off;
    const long long per = (nb + VG_TREE - 1) / VG_TREE;
    const long long b0 = threadIdx.x * per;
    long long local = 0;
    for (long long b = b0; b < b0 + per && b < nb; ++b) local += bsum[off + b];
    long long total;
    long long acc = vg_block_scan_excl(local, wt, &total);
    for (long long b = b0; b < b0 + per && b < nb; ++b) { boff[off + b] = acc; acc += bsum[off + b]; }
    if (threadIdx.x == 0) {
        if (total > I.s_cap) { atomicOr(err, VG_ERR_CAPACITY); total = I.s_cap; }
        s_item[item] = total;
    }
    __syncthreads();
    if (item == 0 && threadIdx.x == 0) {
        if (advance) *it_p = *it_p + 1;
        counts[0] = 0;
        counts[1] = 0;
    }
}

// Batched counterpart of vg_bigfold_kernel.  Blocks [0, VG_BIGFOLD_BLOCKS):
// cubes spanning more than VG_FOLD_MAX windows, list entries (item << 32) |
// cube, tree fold.  The other blocks: cubes spanning 2..VG_FOLD_MAX windows,
// found at window starts over all items' windows (walking the sampling
// kernel's task list, (item, first window) per task_windows windows): the cube
// holding the first sample of window w crosses from window w - 1 when it
// starts there; its partials are left-folded in window order into JS.
__global__ void __launch_bounds__(256)
vg_bbigfold_kernel(const vg_bitem* __restrict__ items, const int* __restrict__ task_item,
                   const int* __restrict__ task_w0, int n_tasks, int task_windows,
                   const long long* __restrict__ list, const int* __restrict__ count, int cap,
                   const long long* __restrict__ nh_sum, const int* __restrict__ run_cube_all,
                   const long long* __restrict__ s_item, const double2* __restrict__ win_first,
                   const double2* __restrict__ win_last, int ws_shift, double2* __restrict__ JS) {
    __shared__ double red[256];
    __shared__ double part[2 * 2048];
    if (blockIdx.x >= VG_BIGFOLD_BLOCKS) {
        const long long stride = (long long)(gridDim.x - VG_BIGFOLD_BLOCKS) * blockDim.x;
        const long long nwt = (long long)n_tasks * task_windows;
        for (long long g = (long long)(blockIdx.x - VG_BIGFOLD_BLOCKS) * blockDim.x + threadIdx.x; g < nwt;
             g += stride) {
            const long long task = g / task_windows;
            const int item = task_item[task];
            const long long w = task_w0[task] + g % task_windows;
            const long long nwin = (s_item[item] + (1LL << ws_shift) - 1) >> ws_shift;
            if (w < 1 || w >= nwin) continue;
            const vg_bitem I = items[item];
            const long long gw = I.win_off + w;
            const long long* nhs = nh_sum + I.nhs_off;
            const int h = run_cube_all[gw << 6];
            const long long st = nhs[h];
            if ((st >> ws_shift) != w - 1) continue;  // starts at window w, or folded from an earlier one
            const long long wb = (nhs[h + 1] - 1) >> ws_shift;
            if (wb - (w - 1) + 1 > VG_FOLD_MAX) continue;  // tree fold
            double2 v = win_last[gw - 1];
            double j1 = v.x, j2 = v.y;
            for (long long u = w; u <= wb; ++u) {
                v = win_first[I.win_off + u];
                j1 = j1 + v.x;
                j2 = j2 + v.y;
            }
            JS[I.cube_off + h] = make_double2(j1, j2);
        }
        return;
    }
    const int n = *count < cap ? *count : cap;
    for (int i = blockIdx.x; i < n; i += VG_BIGFOLD_BLOCKS) {
        const int item = (int)(list[i] >> 32);
        const int h = (int)(list[i] & 0xffffffffLL);
        const vg_bitem I = items[item];
        const long long st = nh_sum[I.nhs_off + h], en = nh_sum[I.nhs_off + h + 1];
        const long long wa = (st >> ws_shift) + I.win_off, wb = ((en - 1) >> ws_shift) + I.win_off;
        double o1, o2;
        vg_fold_tree(wa, wb - wa + 1, win_first, win_last, part, red, o1, o2);
        if (threadIdx.x == 0) JS[I.cube_off + h] = make_double2(o1, o2);
    }
}

// Batched counterpart of vg_scan_cubes_kernel: block b scans 1024 cubes of
// item blk_item[b] (item ranges are padded to 4 cubes, so a thread's 4 cubes
// are one 16-byte load).  A cube writes its own runs when it holds at most
// VG_OWN_RUNS of them; larger cubes go to big_list (absolute run indices,
// written by vg_big_runs_kernel).  counts[0]: fold list length, counts[1]:
// big list length.
__global__ void __launch_bounds__(256)
vg_bscan_cubes_kernel(const vg_bitem* __restrict__ items, const int* __restrict__ blk_item,
                      const int* __restrict__ nh_all, const long long* __restrict__ boff,
                      const long long* __restrict__ s_item, int k_shift,
                      long long* __restrict__ nh_sum_all, int* __restrict__ run_cube_all,
                      long long* __restrict__ fold_list, int* __restrict__ counts, int fold_cap,
                      long long* __restrict__ big_list, int big_cap) {
    __shared__ long long wt[4];
    __shared__ long long slot0;
    const int item = blk_item[blockIdx.x];
    const vg_bitem I = items[item];
    const int* nh = nh_all + I.cube_off;
    long long* nh_sum = nh_sum_all + I.nhs_off;
    int* run_cube = run_cube_all + I.win_off * 64;
    const long long run_cap = I.wcap * 64;
    const int C = I.C;
    const long long base = (blockIdx.x - I.blk_off) * VG_TREE;
    const long long h0 = base + 4 * threadIdx.x;
    const bool full = h0 + 4 <= C;  // item ranges are padded: 16-byte load, 2 x 16-byte stores
    long long v[4];
    long long own = 0;
    if (full) {
        const int4 q = *reinterpret_cast<const int4*>(nh + h0);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = h0 + j < C ? nh[h0 + j] : 0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) own += v[j];
    long long total;
    const long long S = s_item[item];
    const long long ex = vg_block_scan_excl(own, wt, &total) + boff[blockIdx.x];
    long long e[5];
    e[0] = ex;
#pragma unroll
    for (int j = 0; j < 4; ++j) e[j + 1] = e[j] + v[j];
#pragma unroll
    for (int j = 0; j < 5; ++j) e[j] = e[j] < S ? e[j] : S;
    const long long kmask = (1LL << k_shift) - 1;
    long long r0[4], nr[4];
    long long nchunk = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        r0[j] = (e[j] + kmask) >> k_shift;
        long long r1 = (e[j + 1] + kmask) >> k_shift;
        if (r1 > run_cap) r1 = run_cap;
        nr[j] = h0 + j < C ? r1 - r0[j] : 0;
        if (nr[j] > VG_OWN_RUNS) nchunk += (nr[j] + VG_BIG_CHUNK - 1) / VG_BIG_CHUNK;
    }
    long long nblock;
    long long slot = vg_block_scan_excl(nchunk, wt, &nblock);
    if (nblock > 0) {
        if (threadIdx.x == 0) slot0 = atomicAdd(counts + 1, (int)nblock);
        __syncthreads();
        slot += slot0;
    }
    if (full) {
        longlong2* o = reinterpret_cast<longlong2*>(nh_sum + h0);
        o[0] = make_longlong2(e[0], e[1]);
        o[1] = make_longlong2(e[2], e[3]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const long long h = h0 + j;
        if (h < C) {
            if (!full) nh_sum[h] = e[j];
            if (h == C - 1) nh_sum[C] = S;
            const long long n = nr[j];
            if (n > VG_OWN_RUNS) {
                for (long long c = 0; c * VG_BIG_CHUNK < n; ++c, ++slot) {
                    const long long m = n - c * VG_BIG_CHUNK < VG_BIG_CHUNK ? n - c * VG_BIG_CHUNK : VG_BIG_CHUNK;
                    const long long rc = r0[j] + c * VG_BIG_CHUNK;
                    if (slot < big_cap) {
                        big_list[2 * slot] = I.win_off * 64 + rc;
                        big_list[2 * slot + 1] = (m << 32) | h;
                    } else {  // never with big_cap = vg_big_list_cap(total runs); kept correct anyway
                        for (long long r = rc; r < rc + m; ++r) run_cube[r] = (int)h;
                    }
                }
            } else {
                for (long long r = r0[j]; r < r0[j] + n; ++r) run_cube[r] = (int)h;
            }
            const long long st = e[j], en = e[j + 1];
            if (en > st && ((en - 1) >> (k_shift + 6)) - (st >> (k_shift + 6)) + 1 > VG_FOLD_MAX) {
                const int fs = atomicAdd(counts, 1);
                if (fs < fold_cap) fold_list[fs] = ((long long)item << 32) | (long long)h;
            }
        }
    }
}

// One block per (item, dimension).
__global__ void __launch_bounds__(VG_TREE)
vg_brefine_kernel(const vg_bitem* __restrict__ items, const int* __restrict__ rd_item,
                  const int* __restrict__ rd_dim, int n, double alpha, int init,
                  double* __restrict__ x_all, double* __restrict__ weights,
                  unsigned* __restrict__ counts, double2* __restrict__ map64,
                  float2* __restrict__ map32) {
    __shared__ double t[VG_TREE];
    extern __shared__ double dyn[];
    const vg_bitem I = items[rd_item[blockIdx.x]];
    const int dim = rd_dim[blockIdx.x];
    double* x = x_all + I.xe_off + (long long)dim * (n + 1);
    const long long mo = I.map_off + (long long)dim * n;
    if (init) {
        const double step = 1.0 / n;
        for (int j = threadIdx.x; j <= n; j += blockDim.x) x[j] = j * step;
        for (int k = threadIdx.x; k < n; k += blockDim.x) { weights[mo + k] = 0.0; counts[mo + k] = 0u; }
        __syncthreads();
        vg_write_tables(n, x, map64 + mo, map32 + mo);
        return;
    }
    vg_refine_dim<VG_TREE>(n, alpha, x, weights + mo, counts + mo, vg_refine_ws_at(dyn, n), map64 + mo,
                           map32 + mo, t);
}

hipError_t vg_launch_batch_iteration_steps(hipStream_t st, const vg_bitem* items, int n_items,
                                           long long nblk, const int* blk_item, int n_refine,
                                           const int* rd_item, const int* rd_dim, int n,
                                           double alpha, double beta, int k_shift, int hist_cap,
                                           int* nh, long long* nh_sum, double2* JS,
                                           const double2* win_first, const double2* win_last,
                                           double* dh, double* part, double* dsum, long long* bsum,
                                           long long* boff, long long* s_item, int* run_cube,
                                           int* it_p, double* hist, double* x_all,
                                           double* weights, unsigned* counts, double2* map64,
                                           float2* map32, int* err, int phase,
                                           long long* fold_list, int* fold_count, int fold_cap,
                                           const int* task_item, const int* task_w0, int n_tasks,
                                           int task_windows, long long* big_list, int big_cap) {
    if (phase == 0) {  // initial allocation (the map is initialised by vg_launch_batch_refine)
        hipLaunchKernelGGL(vg_balloc_kernel, dim3((unsigned)nblk), dim3(256), 0, st, items, blk_item, 0,
                           dh, dsum, nh, bsum);
    } else {
        const long long wblocks = ((long long)n_tasks * task_windows + 255) / 256;
        hipLaunchKernelGGL(vg_bbigfold_kernel, dim3((unsigned)(VG_BIGFOLD_BLOCKS + wblocks)), dim3(256), 0, st,
                           items, task_item,
                           task_w0, n_tasks, task_windows, fold_list, fold_count, fold_cap, nh_sum, run_cube,
                           s_item, win_first, win_last, 6 + k_shift, JS);
        // the refinement blocks' workspace, or the results blocks' tree buffer
        size_t lds = sizeof(double) * (size_t)(4LL * n + 1 + n / VG_TREE + 2);
        if (lds < sizeof(double) * VG_NRES * 256) lds = sizeof(double) * VG_NRES * 256;
        hipLaunchKernelGGL(vg_bresults_kernel, dim3((unsigned)(nblk + n_refine)), dim3(256), lds, st, items,
                           blk_item, nh, JS, beta, dh, part, n_refine, rd_item, rd_dim, n, alpha, x_all,
                           weights, counts, map64, map32);
        hipLaunchKernelGGL(vg_bfinalize_kernel, dim3(n_items), dim3(256), 0, st, items, part, s_item,
                           it_p, hist_cap, hist, dsum, err);
        hipLaunchKernelGGL(vg_balloc_kernel, dim3((unsigned)nblk), dim3(256), 0, st, items, blk_item, 1,
                           dh, dsum, nh, bsum);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(vg_bscan_blocks_kernel, dim3(n_items), dim3(VG_TREE), 0, st, items, bsum, boff,
                       s_item, it_p, phase == 0 ? 0 : 1, err, fold_count);
    hipLaunchKernelGGL(vg_bscan_cubes_kernel, dim3((unsigned)nblk), dim3(256), 0, st, items, blk_item, nh,
                       boff, s_item, k_shift, nh_sum, run_cube, fold_list, fold_count, fold_cap, big_list,
                       big_cap);
    hipLaunchKernelGGL(vg_big_runs_kernel, dim3(2048), dim3(256), 0, st, big_list, fold_count, big_cap,
                       run_cube);
    return hipGetLastError();
}

hipError_t vg_launch_batch_refine(hipStream_t st, const vg_bitem* items, int n_refine, const int* rd_item,
                                  const int* rd_dim, int n, double alpha, int init, double* x_all,
                                  double* weights, unsigned* counts, double2* map64, float2* map32) {
    const size_t lds = init ? 0 : sizeof(double) * (size_t)(4LL * n + 1 + n / VG_TREE + 2);
    hipLaunchKernelGGL(vg_brefine_kernel, dim3(n_refine), dim3(init ? 256 : VG_TREE), lds, st, items,
                       rd_item, rd_dim, n, alpha, init, x_all, weights, counts, map64, map32);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Work layout of the backward replay (vg_grad).  The cubes of a recorded
// allocation are sorted by sample count, largest first, stably (ties in cube
// order).  The sort key is a size class: n itself up to 128, then 16 classes
// per octave (sizes within a class differ by at most 1/16 of an octave), 255
// from 31,744 samples on; one 8-bit radix pass sorts by it.  Cube c of the
// sorted order gets L_c lanes, a power of two (1 up to VG_GRAD_T samples, about
// n / VG_GRAD_T beyond, at most 64) that depends only on the class, so L does
// not increase along the order: every cube's lanes start at a multiple of its
// L, and no cube's lanes straddle a wave of 64.  lane_start is the exclusive
// prefix sum of L over the order (lane_start[C] = all lanes); wave_first[w] is
// the first cube of wave w (64 w is always a cube's first lane), and
// wave_first[waves] = C.
// ---------------------------------------------------------------------------
#define VG_GRAD_T 16

__device__ inline unsigned vg_size_class(unsigned n) {
    if (n <= 128u) return n;
    const unsigned o = 31u - (unsigned)__clz((int)n);  // n > 128: o >= 7
    if (o > 14u) return 255u;
    return 128u + 16u * (o - 7u) + ((n >> (o - 4u)) & 15u);
}

// Lanes of the cubes of class k: L = 2^ceil(log2(ceil(n_lo / VG_GRAD_T))), at
// most 64, for n_lo the smallest size of the class.
__device__ inline int vg_class_lanes(unsigned k) {
    unsigned nlo = k;
    if (k > 128u) {
        const unsigned o = 7u + (k - 128u) / 16u, f = (k - 128u) % 16u;
        nlo = (16u + f) << (o - 4u);
    }
    const unsigned q = (nlo + VG_GRAD_T - 1) / VG_GRAD_T;
    unsigned L = 1;
    while (L < q && L < 64u) L <<= 1;
    return (int)L;
}

__global__ void __launch_bounds__(256)
vg_grad_keys_kernel(int C, const int* __restrict__ nh, unsigned char* __restrict__ key, int* __restrict__ idx) {
    for (long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x; h < C;
         h += (long long)gridDim.x * blockDim.x) {
        key[h] = (unsigned char)vg_size_class((unsigned)nh[h]);
        idx[h] = (int)h;
    }
}

__global__ void __launch_bounds__(256)
vg_grad_lanes_kernel(int C, const unsigned char* __restrict__ key_sorted, int* __restrict__ lanes) {
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c <= C;
         c += (long long)gridDim.x * blockDim.x)
        lanes[c] = c < C ? vg_class_lanes(key_sorted[c]) : 0;
}

__global__ void __launch_bounds__(256)
vg_grad_waves_kernel(int C, const int* __restrict__ lane_start, int* __restrict__ wave_first) {
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < C;
         c += (long long)gridDim.x * blockDim.x) {
        const int s = lane_start[c];
        if ((s & 63) == 0) wave_first[s >> 6] = (int)c;
        if (c == C - 1) wave_first[(lane_start[C] + 63) >> 6] = C;
    }
}

size_t vg_grad_order_tmp_bytes(int C) {
    size_t sort_bytes = 0, scan_bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, sort_bytes, (const unsigned char*)nullptr,
                                                     (unsigned char*)nullptr, (const int*)nullptr,
                                                     (int*)nullptr, C, 0, 8) != hipSuccess)
        return 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const int*)nullptr, (int*)nullptr, C + 1) !=
        hipSuccess)
        return 0;
    return sort_bytes > scan_bytes ? sort_bytes : scan_bytes;
}

hipError_t vg_launch_grad_order(hipStream_t st, int C, const int* nh, unsigned char* keys,
                                unsigned char* keys_sorted, int* idx, int* order, int* lanes, int* lane_start,
                                int* wave_first, void* tmp, size_t tmp_bytes) {
    const long long blocks = ((long long)C + 256) / 256;
    const unsigned grid = (unsigned)(blocks < 4096 ? blocks : 4096);
    hipLaunchKernelGGL(vg_grad_keys_kernel, dim3(grid), dim3(256), 0, st, C, nh, keys, idx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = tmp_bytes;
    e = hipcub::DeviceRadixSort::SortPairsDescending(tmp, bytes, (const unsigned char*)keys, keys_sorted,
                                                     (const int*)idx, order, C, 0, 8, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(vg_grad_lanes_kernel, dim3(grid), dim3(256), 0, st, C, keys_sorted, lanes);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    bytes = tmp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, (const int*)lanes, lane_start, C + 1, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(vg_grad_waves_kernel, dim3(grid), dim3(256), 0, st, C, lane_start, wave_first);
    return hipGetLastError();
}

long long vg_grad_max_waves(long long C, long long S) {
    // L = 1 for cubes of at most VG_GRAD_T samples, below 4 n / VG_GRAD_T beyond
    const long long lanes = C + (4 * S + VG_GRAD_T - 1) / VG_GRAD_T;
    return (lanes + 63) / 64;
}
