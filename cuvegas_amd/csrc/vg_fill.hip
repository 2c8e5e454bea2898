// VEGAS+ sampling kernel, compiled at run time by hipRTC together with the
// user's integrand (replaces the PTX link of cuda_kernel.cu with the numba
// integrand, reference cuda_wrapper.cu:221-286 and python_wrapper.py:49-59).
//
// One launch evaluates every sample of one iteration (reference vegasFill,
// cuda_kernel.cu:14-76) and produces
//   * per-cube sums of J f and (J f)^2 in a fixed order, without atomics,
//     stored interleaved (one 16-byte store per cube):
//       level 0  each lane walks a run of VG_K consecutive samples (the
//                sample -> cube map comes from the nh prefix sums, so there is
//                no evals[] array) and sums each cube's samples in order;
//       level 1  the 64 runs of a window (one wave) combine the pieces of a
//                cube that crosses runs with a segmented scan of DPP moves
//                (vg_seg_scan64);
//       level 2  a cube that crosses windows leaves its window partials in
//                win_first / win_last; the results kernel folds them;
//   * per-dimension bin sums of (J f)^2 and hit counts, privatised in LDS and
//     merged into global memory with one atomic per touched bin per block
//     (the reference does 2d global atomics per sample, cuda_kernel.cu:60-65).
//
// Compile-time shape (hipRTC -D options): VG_NDIM, VG_NBINS, VG_NSTRAT, VG_K,
// VG_BLOCK, VG_FP32 (0: FP64 sampling and evaluation; 1: FP32 sampling and
// evaluation with FP64 accumulation), VG_LDS_BINS, VG_BIN_DIMS (dimensions
// whose bins one pass keeps in LDS), VG_BIN_COPIES and VG_COPY_DIM0 (LDS bin
// copies of the dimensions from VG_COPY_DIM0 on), VG_LDS_MAP, VG_HAS_DOMAIN,
// VG_PAIR (two samples per lane step), VG_PIPE (one sample per step, the next
// sample's map entries requested before the current one is evaluated).
#include "vg_philox.h"
#include "vg_math.h"

#ifndef VG_PAIR
#define VG_PAIR 1
#endif
#ifndef VG_PIPE
#define VG_PIPE 0
#endif
#ifndef VG_BIN_COPIES
#define VG_BIN_COPIES 1
#endif
#ifndef VG_COPY_DIM0
#define VG_COPY_DIM0 0
#endif

#if VG_FP32
typedef float vg_real;
#else
typedef double vg_real;
#endif

// exp in the working precision, for integrands.  FP64: vg_exp, the project's
// Cody-Waite + Taylor exp that the CPU oracle restates step for step, so FP64
// integrand values (and with them d_h and the integer allocation) are the same
// on the GPU and the CPU.
// FP32: e^x = 2^(x log2 e) with the product carried as hi + lo (the fma gives
// the rounding error of hi exactly, plus x times the tail of log2 e), then
// 2^(hi + lo) = 2^hi (1 + lo ln 2) to second order in lo; v_exp_f32 gives 2^hi
// to 1 ulp, so the result is within about 2 ulp.  The argument is clamped to
// [-150, 88.7]: below, e^x underflows to 0; 88.72 is the overflow threshold.
__device__ inline double vg_expr(double x) { return vg_exp(x); }
__device__ inline float vg_expr(float x) {
    x = __builtin_amdgcn_fmed3f(x, -150.0f, 88.7f);
    const float l2e_hi = 0x1.715476p+0f, l2e_lo = 0x1.4ae0c0p-26f;
    const float hi = x * l2e_hi;
    const float lo = __builtin_fmaf(x, l2e_lo, __builtin_fmaf(x, l2e_hi, -hi));
    const float e = __builtin_amdgcn_exp2f(hi);
    return __builtin_fmaf(e, lo * 0x1.62e430p-1f, e);
}

// Segmented inclusive scan across the 64 lanes of a wave, for keys that do
// not decrease with the lane (each key's lanes are contiguous): s1, s2 become
// the sums of the lane's segment up to and including itself.  DPP moves only
// (VALU; no LDS traffic behind the bin atomics, no waits on them):
//   1-4: row_shr 1, 2, 4, 8 inside each row of 16 lanes (lane l takes lane
//        l - off of its row when that lane has the same key);
//   5:   row_bcast 15, rows 1 and 3: lane 15 of the row before;
//   6:   row_bcast 31, rows 2 and 3: lane 31.
// After step 4 a lane holds its segment's sum within its row; step 5 adds the
// previous row's part to rows 1 and 3, and then lane 31 holds its segment's
// sum over rows 0-1, which step 6 adds to rows 2 and 3.  Each addition is
// (earlier lanes) + (this lane), the order the oracle restates.
template <int CTRL, int ROWS>
__device__ inline double vg_dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int CTRL, int ROWS>
__device__ inline void vg_seg_step(int key, double& s1, double& s2) {
    // lanes the step does not write keep -2, which no key equals (keys >= -1)
    const int kp = __builtin_amdgcn_update_dpp(-2, key, CTRL, ROWS, 0xf, false);
    const double p1 = vg_dpp_f64<CTRL, ROWS>(s1);
    const double p2 = vg_dpp_f64<CTRL, ROWS>(s2);
    if (kp == key) {
        s1 = p1 + s1;
        s2 = p2 + s2;
    }
}
__device__ inline void vg_seg_scan64(int key, double& s1, double& s2) {
    vg_seg_step<0x111, 0xf>(key, s1, s2);  // row_shr:1
    vg_seg_step<0x112, 0xf>(key, s1, s2);  // row_shr:2
    vg_seg_step<0x114, 0xf>(key, s1, s2);  // row_shr:4
    vg_seg_step<0x118, 0xf>(key, s1, s2);  // row_shr:8
    vg_seg_step<0x142, 0xa>(key, s1, s2);  // row_bcast:15, rows 1 and 3
    vg_seg_step<0x143, 0xc>(key, s1, s2);  // row_bcast:31, rows 2 and 3
}
// The value of lane l - 1 (wave_shr:1; lane 0 gets 0).
__device__ inline double vg_lane_before(double v) { return vg_dpp_f64<0x138, 0xf>(v); }

#include "vg_user.h"

#ifndef VG_BATCH
#define VG_BATCH 0
#endif

#if !VG_BATCH
namespace vg_detail {
// The integrand may be written in any of three forms:
//   vg_real integrand(vg_real* x, const vg_real* theta)   parametrised
//   vg_real integrand(vg_real* x)                          config.cuh style
//   int integrand(vg_real* ret, vg_real* x)                extension ABI
template <typename R>
__device__ inline auto call(R* x, const R* th, int, int) -> decltype(integrand(x, th)) {
    return integrand(x, th);
}
template <typename R>
__device__ inline auto call(R* x, const R*, int, long) -> decltype(integrand(x)) {
    return integrand(x);
}
template <typename R>
__device__ inline R call(R* x, const R*, long, long) {
    R r;
    integrand(&r, x);
    return r;
}
}  // namespace vg_detail

constexpr int D = VG_NDIM;
constexpr int NB = VG_NBINS;
constexpr unsigned NS = VG_NSTRAT;
// The sampling tables hold (edge, NB * width) when NB is a power of two: the
// Jacobian factor is read instead of formed, and the in-bin offset carries
// the 1/NB.  Both scalings are exact, so the point edge + width * offset and
// the Jacobian are bit-identical to the unscaled form.
constexpr bool VG_WSCALED = (NB & (NB - 1)) == 0;
constexpr float VG_OFF_SCALE_F = VG_WSCALED ? 1.0f / (float)NB : 1.0f;
constexpr double VG_OFF_SCALE = VG_WSCALED ? 1.0 / (double)NB : 1.0;
__device__ inline vg_real vg_jac_factor(vg_real w) { return VG_WSCALED ? w : (vg_real)NB * w; }
constexpr int K = VG_K;
constexpr int NG = (D + 3) / 4;
constexpr int WPB = VG_BLOCK / 64;
constexpr long long WS = 64LL * K;
constexpr long long vg_ipow(long long b, int e) { return e == 0 ? 1 : b * vg_ipow(b, e - 1); }
constexpr long long NCUBES = vg_ipow(NS, D);

// Pipelined loop with the run's cube boundaries preloaded (VG_RUN_BOUNDS): a
// run of K <= 16 samples crosses at most K / 2 cube boundaries (every cube
// holds at least 2 samples), so they are read once, before the loop, as
// offsets from the run start (one byte each, clamped to K + 1), and the sums
// of the cubes a chunk of VG_RB_CHUNK steps closes (at most half as many)
// wait in registers until the chunk ends.  The loop then issues no global access but the map
// gathers: loads and stores share one counter, so a boundary load or a sum
// store inside the loop makes every wait for a gather wait for it as well.
// On by default in the FP32 form: 4-D Gaussian (1e8 samples, 24 M cubes of
// about 5 samples) fill 0.789 -> 0.756 ms, 8-D double Gaussian unchanged
// within 1% (same-box runs).  The FP64 form, held at 168 registers by its
// launch bound, ran 2% slower at the 8-D shape with it and keeps the loads
// and stores in the loop.  The chunk is the whole run by default: the sums of
// a run's cubes then leave in one burst at its end, so the 128-byte lines of
// per-cube sums are written whole instead of over the run's lifetime (the
// fill's written bytes at the 8-D headline 1.54x -> 1.11x the per-cube sums,
// 4-D Gaussian 1.67x -> 1.05x, against chunks of 2 steps) at the same speed.
#ifndef VG_RUN_BOUNDS
#define VG_RUN_BOUNDS (VG_PIPE && !VG_PAIR && VG_K <= 16 && VG_FP32)
#endif
#if VG_RUN_BOUNDS && !(VG_PIPE && !VG_PAIR && VG_K <= 16)
#error "VG_RUN_BOUNDS needs the pipelined loop (pipe=1) and runs of at most 16 samples"
#endif
#ifndef VG_RB_CHUNK
#define VG_RB_CHUNK 16
#endif
constexpr int RB_N = K >= 2 ? K / 2 : 1;                         // boundaries a run can reach
constexpr int RB_CHUNK = K < VG_RB_CHUNK ? K : VG_RB_CHUNK;      // steps per chunk
constexpr int RB_DEPTH = (RB_CHUNK + 1) / 2;                     // sums a chunk can close

struct vg_sample {
    double jf, jf2;
    unsigned ids[D];
    vg_real x[D];
};

#if VG_FP32
typedef float2 vg_map_t;
#else
typedef double2 vg_map_t;
#endif

// Random words of global sample s of iteration it (all dimension groups).
__device__ inline void vg_words(unsigned (&words)[4 * NG], long long s, unsigned it, unsigned seed,
                                unsigned stream) {
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const vg_u4 v = vg_sample_words((unsigned long long)s, it, (unsigned)g, seed, stream);
        words[4 * g + 0] = v.x;
        words[4 * g + 1] = v.y;
        words[4 * g + 2] = v.z;
        words[4 * g + 3] = v.w;
    }
}

// A sample located in the map: bin ids, in-bin offsets and the (edge, width)
// entries of its bins.  Splitting location from evaluation lets the sampling
// loop request the map entries of the next sample before it evaluates the
// current one.
template <typename MapT>
struct vg_pre {
    unsigned ids[D];
    vg_real frac[D];
    MapT m[D];
};

// Bins and offsets of a sample of cube h from its random words; map is the
// packed (edge, width) table [D][NB] in the working precision.
template <typename MapT>
__device__ inline void vg_locate(vg_pre<MapT>& p, const unsigned (&words)[4 * NG], int h,
                                 const MapT* __restrict__ map) {
    unsigned tmp = (unsigned)h;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const unsigned q = tmp / NS;
        const unsigned k = tmp - q * NS;
        tmp = q;
        const vg_binpos b = vg_bin_of(words[i], k * (unsigned)NB, (unsigned)NB, NS);
#if VG_FP32
        constexpr float inv_ns = 1.0f / (float)NS;
#if VG_NSTRAT <= 256
        // rem * 2^24 + (low >> 9) < 2^32: one conversion rounds the same real
        // number as (float)rem + (float)(low >> 9) * 2^-24 does, times 2^24.
        p.frac[i] = (float)((b.rem << 24) | (unsigned)(b.low >> 9)) * (0x1p-24f * inv_ns * VG_OFF_SCALE_F);
#else
        p.frac[i] = ((float)b.rem + (float)(unsigned)(b.low >> 9) * 0x1p-24f) * (inv_ns * VG_OFF_SCALE_F);
#endif
#else
        constexpr double inv_ns = 1.0 / (double)NS;
#if VG_NSTRAT <= 256
        // rem * 2^33 + low < 2^41 is exact in FP64 (as is rem + low * 2^-33),
        // so scaling it by 2^-33 / n_strat rounds exactly as the sum does.
        const unsigned hi = (b.rem << 1) | (unsigned)(b.low >> 32);
        p.frac[i] = fma((double)hi, 0x1p32, (double)(unsigned)b.low) * (0x1p-33 * inv_ns * VG_OFF_SCALE);
#else
        p.frac[i] = ((double)b.rem + (double)b.low * 0x1p-33) * (inv_ns * VG_OFF_SCALE);
#endif
#endif
        // row pointer of dimension i (uniform: the load takes a scalar base
        // and a 32-bit lane offset instead of a 64-bit address per lane)
        const MapT* __restrict__ row = map + i * NB;
        p.m[i] = row[b.id];
        p.ids[i] = b.id;
    }
}

// Mapped point, Jacobian and J f of a located sample.
template <typename MapT>
__device__ inline void vg_finish(vg_sample& o, const vg_pre<MapT>& p,
                                 const double* __restrict__ domain, double domain_vol,
                                 const vg_real* __restrict__ theta) {
#pragma unroll
    for (int i = 0; i < D; ++i) {
#if VG_FP32
        o.x[i] = fmaf(p.m[i].y, p.frac[i], p.m[i].x);
#else
        o.x[i] = fma(p.m[i].y, p.frac[i], p.m[i].x);
#endif
        o.ids[i] = p.ids[i];
    }
    vg_real jac = (vg_real)1;
#pragma unroll
    for (int i = 0; i < D; ++i) jac = jac * vg_jac_factor(p.m[i].y);
    vg_real xe[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
#if VG_HAS_DOMAIN
        xe[i] = o.x[i] * (vg_real)domain[D + i] + (vg_real)domain[i];
#else
        xe[i] = o.x[i];
#endif
    }
    vg_real f = vg_detail::call<vg_real>(xe, theta, 0, 0);
#if VG_HAS_DOMAIN
    f = (vg_real)domain_vol * f;
#endif
    const vg_real jfr = f * jac;
    o.jf = (double)jfr;
    o.jf2 = o.jf * o.jf;
}

// Evaluate global sample s (cube h) of iteration it.
template <typename MapT>
__device__ inline void vg_eval(vg_sample& o, long long s, int h, unsigned it, unsigned seed,
                               unsigned stream, const MapT* __restrict__ map,
                               const double* __restrict__ domain, double domain_vol,
                               const vg_real* __restrict__ theta) {
    unsigned words[4 * NG];
    vg_words(words, s, it, seed, stream);
    vg_pre<MapT> p;
    vg_locate(p, words, h, map);
    vg_finish(o, p, domain, domain_vol, theta);
}

// Per-block bin partials: FP64, or FP32 when VG_BIN32 is set (FP32 form only).
#if VG_BIN32
typedef float vg_bin_t;
#else
typedef double vg_bin_t;
#endif

extern "C" __global__ void __launch_bounds__(VG_BLOCK)
vg_fill(const vg_map_t* __restrict__ gmap,      // [D][NB] (edge, width)
        const long long* __restrict__ nh_sum,   // [C+1] exclusive prefix of n_h
        const int* __restrict__ run_cube,       // cube holding the first sample of each run
        const long long* __restrict__ s_total,  // samples this iteration
        const int* __restrict__ it_p,           // iteration counter
        double2* __restrict__ JS,                 // [C] per-cube (sum J f, sum (J f)^2)
        double2* __restrict__ win_first, double2* __restrict__ win_last,
        double* __restrict__ weights, unsigned* __restrict__ counts,
        const double* __restrict__ domain, double domain_vol,
        const vg_real* __restrict__ theta, unsigned seed, unsigned stream,
        int bin_lo, int write_sums) {
    // When the bins of all D dimensions do not fit in LDS, the host runs one
    // pass per group of BD dimensions [bin_lo, bin_lo + BD): every pass draws
    // the same samples, adds the bins of its group, and only the pass with
    // write_sums set stores the per-cube sums.
#if VG_LDS_BINS
    constexpr int BD = VG_BIN_DIMS;
    // Dimensions i >= VG_COPY_DIM0 (all bins in one pass only) keep
    // VG_BIN_COPIES copies of their bins, lane l adding into copy l mod
    // VG_BIN_COPIES: in the dimensions whose stratum digit a window's cubes
    // share, all lanes add into the few bins of one stratum, and the copies
    // divide those same-address collisions.  Copy 0 is lw / lc itself; the
    // others follow after them, copy cp at BD NB + (cp - 1) XS (XS: the copy
    // dimensions' bins plus an odd pad, so that one bin's copies fall in
    // different LDS banks), and are added to copy 0 at the flush.
    constexpr int CP = (BD == D && VG_BIN_COPIES > 1 && VG_COPY_DIM0 < D) ? VG_BIN_COPIES : 1;
    constexpr int CD0 = CP > 1 ? VG_COPY_DIM0 : D;
    constexpr int XS = (D - CD0) * NB + 17;
    constexpr int NXB = (CP - 1) * XS;
    __shared__ vg_bin_t lw[BD * NB + NXB];
    __shared__ unsigned lc[BD * NB + NXB];
    for (int i = threadIdx.x; i < BD * NB + NXB; i += VG_BLOCK) { lw[i] = (vg_bin_t)0; lc[i] = 0u; }
#endif
#if VG_LDS_MAP
    __shared__ vg_map_t lmap[D * NB];
    for (int i = threadIdx.x; i < D * NB; i += VG_BLOCK) lmap[i] = gmap[i];
    const vg_map_t* map = lmap;
#else
    const vg_map_t* map = gmap;
#endif
    __syncthreads();

    const long long S = *s_total;
    const unsigned it = (unsigned)*it_p;
    const long long nwin = (S + WS - 1) / WS;
    const int lane = threadIdx.x & 63;
    const long long wave = (long long)blockIdx.x * WPB + (threadIdx.x >> 6);
    const long long nwaves = (long long)gridDim.x * WPB;

    for (long long w = wave; w < nwin; w += nwaves) {
        const long long r = w * 64 + lane;
        const long long rs = r * K;
        const bool active = rs < S;
        const long long re = active ? (rs + K < S ? rs + K : S) : rs;
        int h = 0;
        long long cstart = 0, cend = 0;
#if VG_RUN_BOUNDS
        // bd[]: the next boundaries as offsets from rs, one byte each, next
        // one lowest; bytes past the preloaded ones read 0xff (beyond the run)
        unsigned bd[2] = {~0u, ~0u};
        if (active) {
            h = run_cube[r];
            cstart = nh_sum[h];
            long long e[RB_N];  // all requested before any is used
#pragma unroll
            for (int k = 0; k < RB_N; ++k) {
                const long long hk = (long long)h + 1 + k;
                e[k] = nh_sum[hk < NCUBES ? hk : NCUBES];
            }
#pragma unroll
            for (int k = 0; k < RB_N; ++k) {
                const long long o = e[k] - rs;
                const unsigned b = (unsigned)(o < K + 1 ? o : K + 1);
                bd[k >> 2] = (bd[k >> 2] & ~(0xffu << (8 * (k & 3)))) | (b << (8 * (k & 3)));
            }
        }
#else
        if (active) {
            h = run_cube[r];
            cstart = nh_sum[h];
            cend = nh_sum[h + 1];
        }
#endif
        const long long startF = cstart;
        const bool sb = active && cstart < rs;
        bool hasF = false;
        int hF = 0, npieces = 0;
        double f1 = 0.0, f2 = 0.0, a1 = 0.0, a2 = 0.0;

        // The sample at cend opens cube h + 1: close cube h (or keep it as the
        // run's first piece when it began before the run).
        auto close_cube = [&]() {
            if (npieces == 0 && sb) {
                hasF = true; hF = h; f1 = a1; f2 = a2;
            } else if (write_sums) {
                JS[h] = make_double2(a1, a2);
            }
            ++npieces;
            ++h;
            cstart = cend;
            cend = nh_sum[h + 1];
            a1 = 0.0; a2 = 0.0;
        };
        auto take = [&](const vg_sample& o) {
#pragma unroll
            for (int i = 0; i < D; ++i) {
#if VG_LDS_BINS
                const int li = BD == D ? i : i - bin_lo;  // this pass's slot of dimension i
                if (CP > 1 && i >= CD0) {
                    const int cp = lane % CP;
                    const int a = cp == 0 ? li * NB + (int)o.ids[i]
                                          : BD * NB + (cp - 1) * XS + (i - CD0) * NB + (int)o.ids[i];
                    atomicAdd(&lw[a], (vg_bin_t)o.jf2);
                    atomicAdd(&lc[a], 1u);
                } else if (BD == D || (unsigned)li < (unsigned)BD) {
                    atomicAdd(&lw[li * NB + o.ids[i]], (vg_bin_t)o.jf2);
                    atomicAdd(&lc[li * NB + o.ids[i]], 1u);
                }
#else
                unsafeAtomicAdd(&weights[i * NB + o.ids[i]], o.jf2);
                atomicAdd(&counts[i * NB + o.ids[i]], 1u);
#endif
            }
            a1 = a1 + o.jf;
            a2 = a2 + o.jf2;
        };
#if VG_K >= 2 && VG_PAIR
        // Two samples per step, evaluated independently so that their map
        // lookups and transcendental latencies overlap, then taken in sample
        // order (the per-cube sums are unchanged).  Every cube holds at least
        // 2 samples, so the pair crosses at most one cube boundary and both
        // cubes are known before either evaluation.
        for (int j = 0; j < K; j += 2) {
            const long long s0 = rs + j;
            if (s0 < re) {
                const long long s1 = s0 + 1;
                const bool v1 = s1 < re;
                const bool c0 = s0 == cend;
                const bool c1 = !c0 && v1 && s1 == cend;
                const int h0 = h + (c0 ? 1 : 0);
                const int h1 = h0 + (c1 ? 1 : 0);
                vg_sample o0, o1;
                vg_eval(o0, s0, h0, it, seed, stream, map, domain, domain_vol, theta);
                // past the end of the sample range: evaluate s0 again, unused
                vg_eval(o1, v1 ? s1 : s0, h1, it, seed, stream, map, domain, domain_vol, theta);
                if (c0) close_cube();
                take(o0);
                if (v1) {
                    if (c1) close_cube();
                    take(o1);
                }
            }
        }
#elif VG_PIPE
        // Software-pipelined: the map entries of sample s + 1 are requested
        // (after its Philox words are formed) before sample s is evaluated and
        // its 2 D bin atomics are issued, so the LDS (or cache) latency of the
        // lookups is covered by the evaluation instead of being waited for
        // behind the previous sample's atomics.  Sample s + 1 lies in cube
        // h + (s + 1 == cend), cend being the end of the cube holding s (as
        // updated by close_cube); it is compared only after s has been
        // evaluated, which covers the load of a new cend.  Inactive lanes
        // locate sample rs in cube 0: in bounds, never taken.
        vg_pre<vg_map_t> cur;
        unsigned wn[4 * NG];
        vg_words(wn, rs, it, seed, stream);
        vg_locate(cur, wn, h, map);
#if VG_RUN_BOUNDS
        // cb: offset of the end of cube h; buf[k]: sums of cube h - 1 - k
        // (k < nbuf), closed in this chunk
        unsigned cb = bd[0] & 0xffu;
        double2 buf[RB_DEPTH];
        int nbuf = 0;
#pragma unroll 1
        for (int c0 = 0; c0 < K; c0 += RB_CHUNK) {
#pragma unroll 1
            for (int jj = 0; jj < RB_CHUNK; ++jj) {
                const int j = c0 + jj;
                const long long s = rs + j;
                if (s < re) {
                    if ((unsigned)j == cb) {
                        if (npieces == 0 && sb) {
                            hasF = true; hF = h; f1 = a1; f2 = a2;
                        } else if (write_sums) {
#pragma unroll
                            for (int k = RB_DEPTH - 1; k > 0; --k) buf[k] = buf[k - 1];
                            buf[0] = make_double2(a1, a2);
                            ++nbuf;
                        }
                        ++npieces;
                        ++h;
                        bd[0] = (bd[0] >> 8) | (bd[1] << 24);
                        bd[1] = (bd[1] >> 8) | 0xff000000u;
                        cb = bd[0] & 0xffu;
                        a1 = 0.0; a2 = 0.0;
                    }
                    const bool more = s + 1 < re;
                    if (more) vg_words(wn, s + 1, it, seed, stream);
                    vg_sample o;
                    vg_finish(o, cur, domain, domain_vol, theta);
                    take(o);
                    if (more) vg_locate(cur, wn, h + ((unsigned)(j + 1) == cb ? 1 : 0), map);
                }
            }
#pragma unroll
            for (int k = 0; k < RB_DEPTH; ++k)
                if (k < nbuf) JS[h - 1 - k] = buf[k];
            nbuf = 0;
        }
        // the current cube's end and start as the level-1 step reads them:
        // the end past re exactly when its offset is (offsets clamp at K + 1,
        // and re - rs <= K); a cube opened in the run starts at or after rs
        cend = cb > (unsigned)(re - rs) ? re + 1 : re;
        if (npieces > 0) cstart = rs;
#else
        for (int j = 0; j < K; ++j) {
            const long long s = rs + j;
            if (s < re) {
                if (s == cend) close_cube();
                const bool more = s + 1 < re;
                if (more) vg_words(wn, s + 1, it, seed, stream);
                vg_sample o;
                vg_finish(o, cur, domain, domain_vol, theta);
                take(o);
                if (more) vg_locate(cur, wn, h + (s + 1 == cend ? 1 : 0), map);
            }
        }
#endif
#else
        for (int j = 0; j < K; ++j) {
            const long long s = rs + j;
            if (s < re) {
                if (s == cend) close_cube();
                vg_sample o;
                vg_eval(o, s, h, it, seed, stream, map, domain, domain_vol, theta);
                take(o);
            }
        }
#endif

        // level 1: segmented scan of each run's last piece across the window
        const bool cont = cend > re;
        const int key = active ? h : -1;
        double s1 = active ? a1 : 0.0, s2 = active ? a2 : 0.0;
        vg_seg_scan64(key, s1, s2);
        const double q1 = vg_lane_before(s1), q2 = vg_lane_before(s2);
        const long long wstart = w * WS;
        if (active && write_sums) {
            if (hasF) {
                double t1 = f1, t2 = f2;
                if (lane > 0) { t1 = q1 + t1; t2 = q2 + t2; }
                if (startF < wstart) win_first[w] = make_double2(t1, t2);
                else JS[hF] = make_double2(t1, t2);
            }
            if (!cont) {
                if (cstart < wstart) win_first[w] = make_double2(s1, s2);
                else JS[h] = make_double2(s1, s2);
            } else if (lane == 63) {
                win_last[w] = make_double2(s1, s2);
                if (cstart < wstart) win_first[w] = make_double2(s1, s2);
            }
        }
    }

#if VG_LDS_BINS
    __syncthreads();
    const int nb_pass = (D - bin_lo < BD ? D - bin_lo : BD) * NB;
    for (int i = threadIdx.x; i < nb_pass; i += VG_BLOCK) {
        unsigned c = lc[i];
        vg_bin_t wsum = lw[i];
        if (CP > 1 && i >= CD0 * NB) {  // the other copies of this bin, in copy order
#pragma unroll
            for (int cp = 1; cp < CP; ++cp) {
                const int a = BD * NB + (cp - 1) * XS + (i - CD0 * NB);
                c += lc[a];
                wsum = wsum + lw[a];
            }
        }
        if (c) {
            unsafeAtomicAdd(&weights[bin_lo * NB + i], (double)wsum);
            atomicAdd(&counts[bin_lo * NB + i], c);
        }
    }
#endif
}

// Sample dump: cube, bin ids, mapped point and J f of samples [0, n) of
// iteration it under the current map and allocation (diagnostics and parity
// tests of the sampling step).
extern "C" __global__ void __launch_bounds__(256)
vg_dump(const vg_map_t* __restrict__ gmap, const long long* __restrict__ nh_sum, int C,
        long long n, unsigned it, const double* __restrict__ domain, double domain_vol,
        const vg_real* __restrict__ theta, unsigned seed, unsigned stream,
        long long* __restrict__ cube, int* __restrict__ ids, double* __restrict__ x,
        double* __restrict__ jf) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n || s >= nh_sum[C]) return;
    int lo = 0, hi = C;  // largest h with nh_sum[h] <= s
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        if (nh_sum[mid] <= s) lo = mid; else hi = mid;
    }
    vg_sample o;
    vg_eval(o, s, lo, it, seed, stream, gmap, domain, domain_vol, theta);
    cube[s] = lo;
    for (int i = 0; i < D; ++i) {
        ids[s * D + i] = (int)o.ids[i];
        x[s * D + i] = (double)o.x[i];
    }
    jf[s] = o.jf;
}

#endif  // !VG_BATCH

#if VG_GRAD && !VG_BATCH
// ---------------------------------------------------------------------------
// Backward replay: d I_k / d theta on a frozen iteration (map, allocation and
// Philox counters of forward iteration it).  The integrand supplies
//   vg_real integrand_grad(vg_real* x, const vg_real* theta, vg_real* grad)
// returning f and writing df/dtheta[0..VG_NTHETA).  Samples are regenerated,
// not stored.  Per-cube sums of J df/dtheta and their squares (each cube
// replayed by a block of lanes sized to its sample count, see vg_grad) give
// the stratified variance of the gradient:
//   G_p   = sum_h V/n_h sum_s J df_p
//   Var_p = sum_h (V^2 sum_s (J df_p)^2 / n_h - (V sum_s J df_p / n_h)^2) / (n_h - 1)
// (unbiased per-cube sample variance; n_h >= 2 always)
// Also replays I_k and sigma_k^2 with the forward's expressions (results
// kernel) so value(theta) of the frozen estimator is available for finite
// differences.  part[q][unit] (q < 2 + 2P): fixed-order partials of the
// launch's work units, summed in unit order by vg_grad_reduce.
// ---------------------------------------------------------------------------
constexpr int NT = VG_NTHETA;
// The map is read through L1/L2: the LDS holds the transpose buffers (a
// 48 KB FP32 map staged in LDS as well measured slower).

// Transposed wave sums of NT + 1 values (one value term, NT gradient terms)
// go through a per-wave LDS buffer of 64 rows of VG_TW doubles, in VG_NCH
// column chunks.
constexpr int VG_TW = NT + 1 < 33 ? NT + 1 : 33;
constexpr int VG_NCH = (NT + 1 + VG_TW - 1) / VG_TW;

// J f and J df/dtheta (FP64) of sample s of cube h under the frozen map.
__device__ inline void vg_grad_sample(long long h, long long s, unsigned it, const vg_map_t* __restrict__ gmap,
                                      const double* __restrict__ domain, double domain_vol,
                                      const vg_real* __restrict__ theta, unsigned seed, unsigned stream,
                                      double& jf, double* jg) {
    unsigned words[4 * NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const vg_u4 v = vg_sample_words((unsigned long long)s, it, (unsigned)g, seed, stream);
        words[4 * g + 0] = v.x; words[4 * g + 1] = v.y;
        words[4 * g + 2] = v.z; words[4 * g + 3] = v.w;
    }
    unsigned tmp = (unsigned)h;
    vg_real x[D], dxv[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const unsigned q = tmp / NS;
        const unsigned k = tmp - q * NS;
        tmp = q;
        const vg_binpos b = vg_bin_of(words[i], k * (unsigned)NB, (unsigned)NB, NS);
#if VG_FP32
        constexpr float inv_ns = 1.0f / (float)NS;
        const float frac = ((float)b.rem + (float)(unsigned)(b.low >> 9) * 0x1p-24f) * (inv_ns * VG_OFF_SCALE_F);
        const float2 m = (gmap + i * NB)[b.id];
        x[i] = fmaf(m.y, frac, m.x);
#else
        constexpr double inv_ns = 1.0 / (double)NS;
        const double frac = ((double)b.rem + (double)b.low * 0x1p-33) * (inv_ns * VG_OFF_SCALE);
        const double2 m = (gmap + i * NB)[b.id];
        x[i] = fma(m.y, frac, m.x);
#endif
        dxv[i] = m.y;
    }
    vg_real jac = (vg_real)1;
#pragma unroll
    for (int i = 0; i < D; ++i) jac = jac * vg_jac_factor(dxv[i]);
#if VG_HAS_DOMAIN
#pragma unroll
    for (int i = 0; i < D; ++i) x[i] = x[i] * (vg_real)domain[D + i] + (vg_real)domain[i];
#endif
    vg_real gr[NT];
    vg_real f = integrand_grad(x, theta, gr);
#if VG_HAS_DOMAIN
    f = (vg_real)domain_vol * f;
#pragma unroll
    for (int p = 0; p < NT; ++p) gr[p] = (vg_real)domain_vol * gr[p];
#endif
    jf = (double)(f * jac);
#pragma unroll
    for (int p = 0; p < NT; ++p) jg[p] = (double)(gr[p] * jac);
}

// Sums over the wave's 64 lanes of the NT + 1 values (a0, g[0..NT)), value
// k = 0 being a0 and k = 1 + p being g[p], through the per-wave LDS buffer
// tb[64][VG_TW]: for each chunk c of VG_TW values, every lane writes its row,
// then lane q < VG_TW sums column q (value k = c * VG_TW + q) over the lanes in
// a fixed order (four interleaved partial sums), so the result does not
// depend on timing.  r[c] receives that sum (0 where k > NT or q >= VG_TW).
// Costs VG_TW row writes and 64 column reads per lane and chunk, where a
// butterfly costs 6 shuffle steps per value.
__device__ inline void vg_wave_tsum(double* tb, int lane, double a0, const double* g, double* r) {
    double* row = tb + lane * VG_TW;
#pragma unroll
    for (int c = 0; c < VG_NCH; ++c) {
#pragma unroll
        for (int q = 0; q < VG_TW; ++q) {
            const int k = c * VG_TW + q;
            if (k <= NT) row[q] = k == 0 ? a0 : g[k - 1];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double t = 0.0;
        if (lane < VG_TW && c * VG_TW + lane <= NT) {
            double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
#pragma unroll 4
            for (int l = 0; l < 64; l += 4) {
                t0 = t0 + tb[(l + 0) * VG_TW + lane];
                t1 = t1 + tb[(l + 1) * VG_TW + lane];
                t2 = t2 + tb[(l + 2) * VG_TW + lane];
                t3 = t3 + tb[(l + 3) * VG_TW + lane];
            }
            t = (t0 + t1) + (t2 + t3);
        }
        r[c] = t;
        // the rows are rewritten only after every lane has read its column
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Work layout (built on the device for each recorded allocation, see
// vg_launch_grad_order): the cubes sorted by sample count, largest first;
// sorted cube c owns lanes [lane_start[c], lane_start[c + 1]) of the launch,
// L_c of them (a power of two, about n / 16, 1 for cubes of up to 16
// samples), and wave w covers lanes [64 w, 64 w + 64), which start at cube
// wave_first[w].  Lane j of a cube replays its samples j, j + L, j + 2L, ...,
// so no lane replays more than about 32 samples; the cube's lanes then combine
// their sums by a butterfly over the cube's aligned block of L lanes, the
// first applies the cube expressions, and the wave's contributions are summed
// by two transposed sums into the unit's LDS accumulator (entry 0, 1: value
// terms; 2 + p, 2 + NT + p: gradient terms, each written by one lane), in a
// fixed order.  A work unit (one wave, one block) takes slots_per_unit
// consecutive waves of the layout and writes its accumulator to part[q][unit]
// (value q of 2 + 2 NT, unit-contiguous); vg_grad_reduce sums the units in
// index order.  The largest cubes come first, so the hardware starts the
// longest units first.  The launch bound asks for VG_GRAD_WPS waves per SIMD.
// With 1 the compiler may use up to 512 registers: the FP64 kernel then keeps
// its 286 without scratch spills (one wave per SIMD, the same replay time as
// two waves with 34 spilled), and the FP32 kernel is scheduled into 236 (two
// waves per SIMD still fit), its replay 14% faster than under the bound of 2
// waves (6-D mixture, 32 parameters: 0.78 against 0.90 ms).
#ifndef VG_GRAD_WPS
#define VG_GRAD_WPS 1
#endif
extern "C" __global__ void __launch_bounds__(64, VG_GRAD_WPS)
vg_grad(const vg_map_t* __restrict__ gmap, const int* __restrict__ nh, const long long* __restrict__ nh_sum,
        const int* __restrict__ order, const int* __restrict__ lane_start, const int* __restrict__ wave_first,
        int C, int slots_per_unit, unsigned it, double V, const double* __restrict__ domain, double domain_vol,
        const vg_real* __restrict__ theta, unsigned seed, unsigned stream, double* __restrict__ part) {
    constexpr int NV = 2 + 2 * NT;
    __shared__ double wacc[NV];
    __shared__ double tb[64 * VG_TW];
    __shared__ int lmap[64];
    const vg_map_t* map = gmap;
    const int lane = threadIdx.x;
    for (int q = lane; q < NV; q += 64) wacc[q] = 0.0;
    const long long total = lane_start[C];
    const long long waves = (total + 63) / 64;
    for (int r = 0; r < slots_per_unit; ++r) {
        const long long w = (long long)blockIdx.x * slots_per_unit + r;
        if (w >= waves) break;
        // lane -> sorted cube: the wave's cubes [c0, cend) write their lanes' slots
        const int c0 = wave_first[w], cend = wave_first[w + 1];
        const long long base = 64 * w;
        if (c0 + lane < cend) {
            const int s0 = (int)(lane_start[c0 + lane] - base), s1 = (int)(lane_start[c0 + lane + 1] - base);
            for (int t = s0; t < s1; ++t) lmap[t] = c0 + lane;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const bool act = base + lane < total;
        const int c = act ? lmap[lane] : c0;
        const int ls = lane_start[c];
        const int L = act ? lane_start[c + 1] - ls : 1;
        const int j = act ? (int)(base + lane - ls) : 0;
        const int h = order[c];
        const int n = act ? nh[h] : 0;
        // every lane has read its slot before the next slot's writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double sf = 0.0, sf2 = 0.0, g1[NT], g2[NT];
#pragma unroll
        for (int p = 0; p < NT; ++p) { g1[p] = 0.0; g2[p] = 0.0; }
        const long long s0 = nh_sum[h];
        for (int e = j; e < n; e += L) {
            double jf, jg[NT];
            vg_grad_sample(h, s0 + e, it, map, domain, domain_vol, theta, seed, stream, jf, jg);
            sf = sf + jf;
            sf2 = fma(jf, jf, sf2);
#pragma unroll
            for (int p = 0; p < NT; ++p) { g1[p] = g1[p] + jg[p]; g2[p] = fma(jg[p], jg[p], g2[p]); }
        }
        // The cube's terms split into parts linear in the sums of squares,
        // which each lane forms from its own partial sums,
        //   sigma_h^2 / n_h:  V^2 sum (J f)^2 / n^2        (minus I_h^2 / n)
        //   Var_h,p:          V^2 sum (J df_p)^2 / (n (n-1))  (minus G_h,p^2 / (n-1)),
        // and parts in the squares of the sums, which need the cube's whole
        // sums of J f and J df_p: a butterfly over the cube's aligned block of
        // L lanes (the wave's first cube has the most lanes) gives them to
        // every lane of the block, and the first lane adds I_h, G_h,p and the
        // squared terms.
        const double inv_n = act ? 1.0 / n : 0.0;
        const double q2 = V * V * inv_n;
        sf2 = sf2 * q2 * inv_n;
        const double q2g = act ? q2 / (n - 1) : 0.0;
#pragma unroll
        for (int p = 0; p < NT; ++p) g2[p] = g2[p] * q2g;
        const int lmax = lane_start[c0 + 1] - lane_start[c0];
        for (int o = 1; o < lmax; o <<= 1) {
            const bool in = o < L;
            const double t = __shfl_xor(sf, o, 64);
            sf = in ? sf + t : sf;
#pragma unroll
            for (int p = 0; p < NT; ++p) {
                const double u = __shfl_xor(g1[p], o, 64);
                g1[p] = in ? g1[p] + u : g1[p];
            }
        }
        if (act && j == 0) {
            const double Ih = sf * inv_n * V;
            sf = Ih;
            sf2 = sf2 - Ih * Ih * inv_n;
            const double vn = V * inv_n, in1 = 1.0 / (n - 1);
#pragma unroll
            for (int p = 0; p < NT; ++p) {
                const double Gh = g1[p] * vn;
                g1[p] = Gh;
                g2[p] = g2[p] - Gh * Gh * in1;
            }
        } else {
            sf = 0.0;
#pragma unroll
            for (int p = 0; p < NT; ++p) g1[p] = 0.0;
        }
        // sums of I_h and sigma_h^2 / n_h (k = 0), of G_h,p and Var_h,p (k = 1 + p)
        double c1[VG_NCH], c2[VG_NCH];
        vg_wave_tsum(tb, lane, sf, g1, c1);
        vg_wave_tsum(tb, lane, sf2, g2, c2);
#pragma unroll
        for (int cc = 0; cc < VG_NCH; ++cc) {
            const int kk = cc * VG_TW + lane;
            if (lane >= VG_TW || kk > NT) continue;
            if (kk == 0) {
                wacc[0] = wacc[0] + c1[cc];
                wacc[1] = wacc[1] + c2[cc];
            } else {
                wacc[1 + kk] = wacc[1 + kk] + c1[cc];
                wacc[1 + NT + kk] = wacc[1 + NT + kk] + c2[cc];
            }
        }
    }
    __syncthreads();
    for (int q = lane; q < NV; q += 64) part[(long long)q * gridDim.x + blockIdx.x] = wacc[q];
}

// out[q] = sum over units u < n_units of part[q][u], one block of 1024 threads
// per q: thread t adds units t, t + 1024, ... into four interleaved partial
// sums (independent loads in flight), then a fixed pairwise tree over the
// threads, so the sum does not depend on timing.
extern "C" __global__ void __launch_bounds__(1024)
vg_grad_reduce(const double* __restrict__ part, int n_units, double* __restrict__ out) {
    __shared__ double t[1024];
    const int q = blockIdx.x;
    const double* col = part + (long long)q * n_units;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int u = threadIdx.x;
    for (; u + 3 * 1024 < n_units; u += 4 * 1024) {
        a0 = a0 + col[u];
        a1 = a1 + col[u + 1024];
        a2 = a2 + col[u + 2 * 1024];
        a3 = a3 + col[u + 3 * 1024];
    }
    for (; u < n_units; u += 1024) a0 = a0 + col[u];
    t[threadIdx.x] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    for (int w = 512; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) t[threadIdx.x] = t[threadIdx.x] + t[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = t[0];
}
#endif

#if VG_BATCH
// ---------------------------------------------------------------------------
// Batched sampling: many independent integrals (items), each with its own
// dimension (<= VG_MAXD), strata, map, allocation and parameter block, in one
// launch per iteration.  The integrand has the form
//   vg_real integrand(vg_real* x, const vg_real* theta, int d)
// where theta points at the item's parameters.  Item i draws from Philox
// stream i, so it reproduces a single-integral run with stream_id = i.
//
// Work is a static list of tasks (item, chunk of VG_TASKW * VG_BLOCK/64
// windows); a block takes whole tasks, so its LDS bins belong to one item at a
// time and are merged into the item's global bins at the end of the task.  The
// host launches one block per task by default (the hardware then hands tasks
// to CUs as they free up, which balances tasks of unequal cost) or a resident
// grid that strides over the list.  Windows past the item's sample count this
// iteration are skipped.
// ---------------------------------------------------------------------------
// Item descriptor; the host writes the same 80-byte layout (vg_host.cpp).
struct vg_item {
    long long cube_off;  // first cube of the item (nh, per-cube sums, dh)
    long long nhs_off;   // first entry of the item's nh_sum (C+1 entries)
    long long win_off;   // first window (win_first / win_last); runs start at 64 * win_off
    double V;            // cube volume n_strat^-d
    double inv_ns;       // 1.0 / n_strat
    int d, C, map_off, theta_off;  // dims, cubes, first (dim, bin) entry, first parameter
    unsigned ns, ns_m, ns_sh1, ns_sh2;  // n_strat and its division magic
    float inv_ns_f, pad;
};

// floor(x / d) for the item's n_strat (Granlund & Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, fig. 4.1).
__device__ inline unsigned vg_udiv(unsigned x, unsigned m, unsigned sh1, unsigned sh2) {
    const unsigned t1 = __umulhi(m, x);
    return (t1 + ((x - t1) >> sh1)) >> sh2;
}

constexpr int MD = VG_MAXD;
constexpr int NB = VG_NBINS;
// table convention as in the single-integral kernels: (edge, NB * width)
// when NB is a power of two, the 1/NB carried by the in-bin offset
constexpr bool VG_WSCALED = (NB & (NB - 1)) == 0;
constexpr float VG_OFF_SCALE_F = VG_WSCALED ? 1.0f / (float)NB : 1.0f;
constexpr double VG_OFF_SCALE = VG_WSCALED ? 1.0 / (double)NB : 1.0;
__device__ inline vg_real vg_jac_factor(vg_real w) { return VG_WSCALED ? w : (vg_real)NB * w; }
constexpr int K = VG_K;
constexpr int WPB = VG_BLOCK / 64;
constexpr long long WS = 64LL * K;

#if VG_FP32
typedef float2 vg_map_t;
#else
typedef double2 vg_map_t;
#endif

// Strata digits of cube h as kn[i] = k_i * NB (k_i the i-th base-n_strat digit).
template <int DD>
__device__ inline void vg_bdigits(unsigned* kn, int h, const vg_item& I) {
    unsigned tmp = (unsigned)h;
#pragma unroll
    for (int i = 0; i < DD; ++i) {
        const unsigned q = vg_udiv(tmp, I.ns_m, I.ns_sh1, I.ns_sh2);
        kn[i] = (tmp - q * I.ns) * (unsigned)NB;
        tmp = q;
    }
}

// kn of cube h + 1 from kn of cube h: the digits counted up like an odometer
// (branch-free carry), the same values vg_bdigits gives for h + 1.
template <int DD>
__device__ inline void vg_bdigits_next(unsigned* kn, const vg_item& I) {
    const unsigned top = I.ns * (unsigned)NB;
    unsigned carry = NB;
#pragma unroll
    for (int i = 0; i < DD; ++i) {
        const unsigned v = kn[i] + carry;
        carry = v == top ? (unsigned)NB : 0u;
        kn[i] = v == top ? 0u : v;
    }
}

// J f of sample s of the cube with strata digits kn (vg_bdigits) of an item of
// dimension DD; bin ids to ids[0..DD).
template <int DD>
__device__ inline void vg_beval(double& jf, unsigned* ids, long long s, const unsigned* kn, unsigned it,
                                unsigned seed, unsigned stream, const vg_item& I,
                                const vg_map_t* __restrict__ map, const vg_real* __restrict__ theta) {
    constexpr int GG = (DD + 3) / 4;
    unsigned words[4 * GG];
#pragma unroll
    for (int g = 0; g < GG; ++g) {
        const vg_u4 v = vg_sample_words((unsigned long long)s, it, (unsigned)g, seed, stream);
        words[4 * g + 0] = v.x; words[4 * g + 1] = v.y;
        words[4 * g + 2] = v.z; words[4 * g + 3] = v.w;
    }
    vg_real x[DD], jac = (vg_real)1;
#pragma unroll
    for (int i = 0; i < DD; ++i) {
        const unsigned long long L = ((unsigned long long)words[i] * 2ull + 1ull) * (unsigned long long)NB;
        const unsigned Q = kn[i] + (unsigned)(L >> 33);
        const unsigned id = vg_udiv(Q, I.ns_m, I.ns_sh1, I.ns_sh2);
        const unsigned rem = Q - id * I.ns;
        const unsigned long long low = L & ((1ull << 33) - 1ull);
#if VG_FP32
        const float frac = ((float)rem + (float)(unsigned)(low >> 9) * 0x1p-24f) * (I.inv_ns_f * VG_OFF_SCALE_F);
        const float2 m = (map + i * NB)[id];
        x[i] = fmaf(m.y, frac, m.x);
#else
        const double frac = ((double)rem + (double)low * 0x1p-33) * (I.inv_ns * VG_OFF_SCALE);
        const double2 m = (map + i * NB)[id];
        x[i] = fma(m.y, frac, m.x);
#endif
        jac = jac * vg_jac_factor(m.y);
        ids[i] = id;
    }
    const vg_real f = integrand(x, theta, DD);
    jf = (double)(f * jac);
}

// Copies of the block's LDS bins (VG_BCOPIES): lane l adds into copy l mod
// VG_BCOPIES, copies VG_BXS entries apart (an odd count, so that one bin's
// copies sit in different banks), summed in copy order at the task's flush.
// With 256 bins per dimension the 64 lanes of a wave meet in one bin often
// (and in all of them when a window's cubes share a stratum); the copies
// divide those collisions.
#ifndef VG_BCOPIES
#define VG_BCOPIES 1
#endif
constexpr int VG_BXS = MD * NB + 17;

// Windows per wave per task: a task covers VG_TASKW * (VG_BLOCK / 64) windows.
#ifndef VG_TASKW
#define VG_TASKW 8
#endif

// The windows of one task of an item of dimension DD (the block's LDS bins
// hold that item's partials).
template <int DD>
__device__ inline void vg_btask(const vg_item& I, int item, long long w0, long long S, long long nwin,
                                unsigned it, const vg_map_t* __restrict__ map,
                                const long long* __restrict__ nhs, const int* __restrict__ run_cube,
                                double2* __restrict__ js,
                                double2* __restrict__ win_first, double2* __restrict__ win_last,
                                const vg_real* __restrict__ theta, unsigned seed, double* lw,
                                unsigned* lc) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int cofs = VG_BCOPIES > 1 ? (lane % VG_BCOPIES) * VG_BXS : 0;  // this lane's bin copy
    for (int tw = 0; tw < VG_TASKW; ++tw) {
        const long long w = w0 + wid + (long long)tw * WPB;
        if (w >= nwin) continue;
        const long long r = w * 64 + lane;
        const long long rs = r * K;
        const bool active = rs < S;
        // the run's samples and cube ends as offsets from rs (32-bit; a cube
        // end past the run is clamped to K + 1)
        const int n = active ? (int)(rs + K < S ? K : S - rs) : 0;
        const long long wstart = w * WS;
        int h = 0, ce = 0;
        bool sb = false, f_before = false;  // first cube started before the run / the window
        if (active) {
            h = run_cube[I.win_off * 64 + r];
            const long long cstart = nhs[h], cend = nhs[h + 1];
            ce = (int)(cend - rs < K + 1 ? cend - rs : K + 1);
            sb = cstart < rs;
            f_before = cstart < wstart;
        }
        unsigned kn[DD];
        vg_bdigits<DD>(kn, h, I);
        bool hasF = false;
        int hF = 0, npieces = 0;
        double f1 = 0.0, f2 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int j = 0; j < K; ++j) {
            if (j < n) {
                if (j == ce) {
                    if (npieces == 0 && sb) { hasF = true; hF = h; f1 = a1; f2 = a2; }
                    else js[h] = make_double2(a1, a2);
                    ++npieces;
                    ++h;
                    vg_bdigits_next<DD>(kn, I);
                    const long long cend = nhs[h + 1];
                    ce = (int)(cend - rs < K + 1 ? cend - rs : K + 1);
                    a1 = 0.0; a2 = 0.0;
                }
                double jf;
                unsigned ids[DD];
                vg_beval<DD>(jf, ids, rs + j, kn, it, seed, (unsigned)item, I, map, theta);
                const double jq = jf * jf;
#pragma unroll
                for (int i = 0; i < DD; ++i) {
                    atomicAdd(&lw[cofs + i * NB + ids[i]], jq);
                    atomicAdd(&lc[cofs + i * NB + ids[i]], 1u);
                }
                a1 = a1 + jf;
                a2 = a2 + jq;
            }
        }
        // the current cube runs past the run; it started before the window
        // only if it is the run's first cube and that one did
        const bool cont = ce > n;
        const bool c_before = npieces == 0 && f_before;
        const int key = active ? h : -1;
        double s1 = active ? a1 : 0.0, s2 = active ? a2 : 0.0;
        vg_seg_scan64(key, s1, s2);
        const double q1 = vg_lane_before(s1), q2 = vg_lane_before(s2);
        const long long gw = I.win_off + w;
        if (active) {
            if (hasF) {
                double t1 = f1, t2 = f2;
                if (lane > 0) { t1 = q1 + t1; t2 = q2 + t2; }
                if (f_before) win_first[gw] = make_double2(t1, t2);
                else js[hF] = make_double2(t1, t2);
            }
            if (!cont) {
                if (c_before) win_first[gw] = make_double2(s1, s2);
                else js[h] = make_double2(s1, s2);
            } else if (lane == 63) {
                win_last[gw] = make_double2(s1, s2);
                if (c_before) win_first[gw] = make_double2(s1, s2);
            }
        }
    }
}

// Work is a static list of tasks (item, first window); the task body is
// specialised for the item's dimension (a switch that is uniform over the
// block), so every loop over dimensions unrolls.
extern "C" __global__ void __launch_bounds__(VG_BLOCK)
vg_bfill(const vg_item* __restrict__ items, const int* __restrict__ task_item,
         const int* __restrict__ task_w0, int ntasks, const vg_map_t* __restrict__ gmap,
         const long long* __restrict__ nh_sum, const int* __restrict__ run_cube,
         const long long* __restrict__ s_item, const int* __restrict__ it_p,
         double2* __restrict__ JS,
         double2* __restrict__ win_first, double2* __restrict__ win_last,
         double* __restrict__ weights, unsigned* __restrict__ counts,
         const vg_real* __restrict__ theta_all, unsigned seed) {
    __shared__ double lw[(VG_BCOPIES - 1) * VG_BXS + MD * NB];
    __shared__ unsigned lc[(VG_BCOPIES - 1) * VG_BXS + MD * NB];
    const unsigned it = (unsigned)*it_p;
    for (int task = blockIdx.x; task < ntasks; task += gridDim.x) {
        const int item = task_item[task];
        const long long w0 = task_w0[task];
        const long long S = s_item[item];
        const long long nwin = (S + WS - 1) / WS;
        if (w0 >= nwin) continue;  // uniform over the block
        const vg_item I = items[item];
        const int d = I.d;
        const vg_map_t* map = gmap + I.map_off;
        const long long* nhs = nh_sum + I.nhs_off;
        const vg_real* theta = theta_all + I.theta_off;
        double2* js = JS + I.cube_off;
        for (int i = threadIdx.x; i < d * NB; i += VG_BLOCK) {
#pragma unroll
            for (int cp = 0; cp < VG_BCOPIES; ++cp) { lw[cp * VG_BXS + i] = 0.0; lc[cp * VG_BXS + i] = 0u; }
        }
        __syncthreads();
        switch (d) {
#define VG_BTASK_CASE(n)                                                                         \
    case n:                                                                                      \
        if constexpr (n <= MD)                                                                   \
            vg_btask<n>(I, item, w0, S, nwin, it, map, nhs, run_cube, js, win_first, win_last,   \
                        theta, seed, lw, lc);                                                    \
        break;
            VG_BTASK_CASE(1) VG_BTASK_CASE(2) VG_BTASK_CASE(3) VG_BTASK_CASE(4)
            VG_BTASK_CASE(5) VG_BTASK_CASE(6) VG_BTASK_CASE(7) VG_BTASK_CASE(8)
            VG_BTASK_CASE(9) VG_BTASK_CASE(10) VG_BTASK_CASE(11) VG_BTASK_CASE(12)
            VG_BTASK_CASE(13) VG_BTASK_CASE(14) VG_BTASK_CASE(15) VG_BTASK_CASE(16)
#undef VG_BTASK_CASE
            default: break;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < d * NB; i += VG_BLOCK) {
            unsigned c = lc[i];
            double wsum = lw[i];
#pragma unroll
            for (int cp = 1; cp < VG_BCOPIES; ++cp) {  // the other copies, in copy order
                c += lc[cp * VG_BXS + i];
                wsum = wsum + lw[cp * VG_BXS + i];
            }
            if (c) {
                unsafeAtomicAdd(&weights[I.map_off + i], wsum);
                atomicAdd(&counts[I.map_off + i], c);
            }
        }
        __syncthreads();
    }
}
#endif  // VG_BATCH
