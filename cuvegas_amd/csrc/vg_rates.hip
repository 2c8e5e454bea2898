// Issue costs of the instructions the sampling kernel is made of, measured on
// the current device: the ceilings its roofline figures are read against.  No
// counterpart in the reference (which reports no roofline).
//
// The sampling kernel (vg_fill) spends its time in two places: the VALU
// (Philox, bin arithmetic, map lookup, the integrand) and the LDS, where every
// sample adds its (J f)^2 and a count into one bin per dimension with
// ds_add_f64 and ds_add_u32 at addresses as random as the samples.  Neither
// cost follows from published figures: gfx950 issues some VALU instructions
// at 32 lanes per SIMD cycle and others at 16, and the LDS serialises atomics
// that meet in one bank.  Each kernel below keeps one kind of instruction
// busy and is timed with HIP events:
//   * LDS: blocks of 1024 threads (16 waves on a CU, as in the sampling
//     kernel), each lane adding into 8 fixed slots of a 1024-slot array
//     (random slots, or slots 64 apart by lane: one bank per lane);
//   * VALU: 8 independent dependency chains per lane of one operation, so
//     the issue rate, not the latency, sets the time.
// Costs are cycles of the device's nominal clock per wave64 instruction: CU
// cycles for the LDS (one LDS per CU), SIMD cycles for the VALU (4 SIMDs per
// CU).
#include <hip/hip_runtime.h>

#include "../../include/cuvegas.h"
#include "vg_internal.h"

namespace {

constexpr int kSlots = 1024;
constexpr int kLdsIters = 256;
constexpr int kValuIters = 1024;

__device__ inline unsigned vg_rate_hash(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// MODE 0: ds_add_f64, 1: ds_add_u32, 2: both per slot (the sampling kernel's
// pair).  RANDOM: slots drawn by a hash of (thread, block, j); otherwise slot
// (wave * 8 + j) * 64 + lane, 64 consecutive slots per instruction.
template <int MODE, bool RANDOM>
__global__ void __launch_bounds__(1024) vg_rate_lds_kernel(unsigned* out, unsigned seed) {
    __shared__ double lw[kSlots];
    __shared__ unsigned lc[kSlots];
    for (int i = threadIdx.x; i < kSlots; i += blockDim.x) {
        lw[i] = 0.0;
        lc[i] = 0u;
    }
    __syncthreads();
    const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    unsigned a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
        a[j] = RANDOM ? vg_rate_hash(threadIdx.x * 8u + (unsigned)j + seed * 131071u + blockIdx.x * 7919u) % kSlots
                      : ((wave * 8u + (unsigned)j) * 64u + lane) % kSlots;
    const double v = 1.0 + threadIdx.x * 1e-9;
    for (int i = 0; i < kLdsIters; ++i) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (MODE != 1) atomicAdd(&lw[a[j]], v);
            if (MODE != 0) atomicAdd(&lc[a[j]], 1u);
        }
    }
    __syncthreads();
    const unsigned k = threadIdx.x % kSlots;
    out[blockIdx.x * blockDim.x + threadIdx.x] = lc[k] + (unsigned)lw[k];
}

// One VALU operation per step on 8 independent chains:
//   0 v_fma_f32 (fmamk), 1 v_fma_f64, 2 v_add_u32 + v_lshrrev_b32 (two
//   operations), 3 v_mad_u64_u32 (64-bit accumulate of a 32x32 product).
template <int OP>
__global__ void __launch_bounds__(256) vg_rate_valu_kernel(unsigned* out, unsigned seed) {
    float f[8];
    double g[8];
    unsigned u[8];
    unsigned long long m[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        f[c] = (float)(threadIdx.x + c + seed) * 1e-3f;
        g[c] = (double)(threadIdx.x + c + seed) * 1e-3;
        u[c] = threadIdx.x * 7u + (unsigned)c + seed;
        m[c] = u[c];
    }
    for (int i = 0; i < kValuIters; ++i) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (OP == 0) f[c] = __builtin_fmaf(f[c], 0.999f, 1e-4f);
            if (OP == 1) g[c] = __builtin_fma(g[c], 0.999, 1e-4);
            if (OP == 2) u[c] = u[c] + (u[c] >> 1);
            if (OP == 3) m[c] = (unsigned long long)(unsigned)m[c] * 0xD2511F53ull + m[c];
        }
    }
    unsigned r = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        if (OP == 0) r ^= __float_as_uint(f[c]);
        if (OP == 1) r ^= (unsigned)__double_as_longlong(g[c]);
        if (OP == 2) r ^= u[c];
        if (OP == 3) r ^= (unsigned)m[c] ^ (unsigned)(m[c] >> 32);
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

typedef void (*vg_rate_fn)(unsigned*, unsigned);

struct vg_rate_case {
    vg_rate_fn fn;
    int block;
    int blocks_per_cu;
    int iters;
    int per_step;  // instructions per slot (LDS) or per chain step (VALU)
    bool lds;
};

}  // namespace

extern "C" int cvg_device_rates(double* cycles, int n) {
    if (!cycles || n < CVG_RATE_COUNT) return vg_fail(CVG_ERR_INVALID, "cvg_device_rates needs CVG_RATE_COUNT outputs");
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return vg_fail(CVG_ERR_NODEVICE, "no HIP device visible");
    const double clock_hz = prop.clockRate * 1e3;
    const int n_cu = prop.multiProcessorCount;
    const vg_rate_case cases[CVG_RATE_COUNT] = {
        {vg_rate_lds_kernel<0, true>, 1024, 4, kLdsIters, 1, true},
        {vg_rate_lds_kernel<1, true>, 1024, 4, kLdsIters, 1, true},
        {vg_rate_lds_kernel<2, true>, 1024, 4, kLdsIters, 2, true},
        {vg_rate_lds_kernel<0, false>, 1024, 4, kLdsIters, 1, true},
        {vg_rate_lds_kernel<1, false>, 1024, 4, kLdsIters, 1, true},
        {vg_rate_valu_kernel<0>, 256, 16, kValuIters, 1, false},
        {vg_rate_valu_kernel<1>, 256, 16, kValuIters, 1, false},
        {vg_rate_valu_kernel<2>, 256, 16, kValuIters, 2, false},
        {vg_rate_valu_kernel<3>, 256, 16, kValuIters, 1, false},
    };
    const long long max_threads = (long long)n_cu * 16 * 256;  // the largest grid below
    unsigned* out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = CVG_OK;
    if (hipMalloc((void**)&out, (size_t)max_threads * sizeof(unsigned)) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = vg_fail(CVG_ERR_HIP, "cvg_device_rates: allocation of the probe buffers failed");
    }
    for (int c = 0; c < CVG_RATE_COUNT && rc == CVG_OK; ++c) {
        const vg_rate_case& k = cases[c];
        const unsigned grid = (unsigned)(n_cu * k.blocks_per_cu);
        float best = 1e30f;
        for (int rep = 0; rep < 3 && rc == CVG_OK; ++rep) {  // the first launch also warms up
            float ms = 0.f;
            hipLaunchKernelGGL(k.fn, dim3(grid), dim3(k.block), 0, st, out, 1u + (unsigned)rep);
            if (hipEventRecord(e0, st) != hipSuccess) rc = vg_fail(CVG_ERR_HIP, "cvg_device_rates: event record failed");
            hipLaunchKernelGGL(k.fn, dim3(grid), dim3(k.block), 0, st, out, 7u + (unsigned)rep);
            if (rc == CVG_OK && (hipGetLastError() != hipSuccess || hipEventRecord(e1, st) != hipSuccess ||
                                 hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = vg_fail(CVG_ERR_HIP, "cvg_device_rates: probe launch failed");
            if (rc == CVG_OK && ms < best) best = ms;
        }
        if (rc != CVG_OK) break;
        const double waves = (double)grid * (k.block / 64);
        const double instr = waves * k.iters * 8.0 * k.per_step;     // wave64 instructions in the launch
        const double units = k.lds ? n_cu : 4.0 * n_cu;               // LDS units or SIMDs
        cycles[c] = best * 1e-3 * clock_hz * units / instr;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (out) (void)hipFree(out);
    return rc;
}
