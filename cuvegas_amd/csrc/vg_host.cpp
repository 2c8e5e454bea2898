// Host driver and C ABI of the MI355X VEGAS+ integrator.
//
// Owns the hipRTC compile of the sampling kernel, the device state of one
// integral and the device-resident iteration loop.  Mirrors the reference's
// host driver (vegas_cuda/src/vegas_cuda_extension/cuda_wrapper.cu:290-872)
// without its per-iteration host synchronisation: one iteration is seven
// kernels on one stream, optionally captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cuvegas.h"
#include "vg_steps.h"
#include "vg_internal.h"

#include "vg_embed.inc"  // kVgFillSrc, kVgPhiloxSrc, kVgMathSrc (generated from vg_fill.hip, vg_philox.h, vg_math.h)

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define VG_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(CVG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

constexpr int kLdsBytes = 160 * 1024;  // gfx950 LDS per workgroup
constexpr int kFillBlock = 1024;
// Backward replay: the most work units (waves) one launch writes partials for.
constexpr long long kGradMaxUnits = 65536;
constexpr int kDefaultRun = 16;

bool checked_cubes(int ns, int d, long long* C) {
    long long r = 1;
    for (int i = 0; i < d; ++i) {
        r *= ns;
        if (r > 2147483647LL) return false;
    }
    *C = r;
    return true;
}

int validate(const cvg_params* p) {
    if (!p) return fail(CVG_ERR_INVALID, "params is NULL");
    if (p->n_dim < 1 || p->n_dim > 64) return fail(CVG_ERR_INVALID, "n_dim must be in 1..64");
    if (p->n_intervals < 2) return fail(CVG_ERR_INVALID, "n_intervals must be >= 2");
    if (p->n_strat < 1) return fail(CVG_ERR_INVALID, "n_strat must be >= 1");
    if ((unsigned long long)p->n_strat * (unsigned long long)p->n_intervals >= (1ull << 32))
        return fail(CVG_ERR_INVALID, "n_strat * n_intervals must be < 2^32");
    if (p->n_strat >= (1 << 24)) return fail(CVG_ERR_INVALID, "n_strat must be < 2^24");
    long long C;
    if (!checked_cubes(p->n_strat, p->n_dim, &C))
        return fail(CVG_ERR_INVALID, "n_strat^n_dim must be < 2^31");
    if (p->precision != CVG_FP64 && p->precision != CVG_FP32C)
        return fail(CVG_ERR_INVALID, "precision must be CVG_FP64 or CVG_FP32C");
    if (p->run_len < 0 || p->run_len > 4096 || (p->run_len & (p->run_len - 1)))
        return fail(CVG_ERR_INVALID, "run_len must be 0 or a power of two up to 4096");
    if (p->has_domain && !p->domain) return fail(CVG_ERR_INVALID, "has_domain set but domain is NULL");
    if (p->n_theta < 0) return fail(CVG_ERR_INVALID, "n_theta must be >= 0");
    return CVG_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Program: the hipRTC-compiled sampling kernel for one (integrand, shape).
// ---------------------------------------------------------------------------
struct cvg_program {
    int d = 0, nbins = 0, nstrat = 0, precision = 0, K = kDefaultRun, has_domain = 0;
    int block = kFillBlock, lds_bins = 0, lds_map = 0, bin_dims = 1, bin_copies = 1, copy_dim0 = 0;
    int has_grad = 0, n_theta = 0, bin32 = 0, pair = 1, pipe = 0;
    int batch = 0, max_dim = 0, task_w = 8, fill_grid = 1;
    std::string log;
    std::vector<char> code;
    double compile_ms = 0.0;
    // module loaded on first use, on the device current at that time
    int device = -1;
    hipModule_t mod = nullptr;
    hipFunction_t fill = nullptr, dump = nullptr, grad = nullptr, grad_reduce = nullptr, bfill = nullptr;
    int fill_blocks_per_cu = 0;
    // guards the module fields: states created from several threads over one
    // program load it once
    std::mutex load_mu;

    int load() {
        int dev;
        VG_HIP(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lock(load_mu);
        if (mod && dev == device) return CVG_OK;
        if (mod) return fail(CVG_ERR_INVALID, "program was loaded on another device");
        const double t0 = now_ms();
        VG_HIP(hipModuleLoadData(&mod, code.data()));
        const int rc = bind_functions();
        if (rc) {
            // leave the program unloaded, as before the call
            (void)hipModuleUnload(mod);
            mod = nullptr;
            fill = dump = grad = grad_reduce = bfill = nullptr;
            return rc;
        }
        if (fill_blocks_per_cu < 1) fill_blocks_per_cu = 1;
        device = dev;
        compile_ms += now_ms() - t0;
        return CVG_OK;
    }
    int bind_functions() {
        if (batch) {
            VG_HIP(hipModuleGetFunction(&bfill, mod, "vg_bfill"));
            VG_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&fill_blocks_per_cu, bfill, block, 0));
        } else {
            VG_HIP(hipModuleGetFunction(&fill, mod, "vg_fill"));
            VG_HIP(hipModuleGetFunction(&dump, mod, "vg_dump"));
            if (has_grad) {
                VG_HIP(hipModuleGetFunction(&grad, mod, "vg_grad"));
                VG_HIP(hipModuleGetFunction(&grad_reduce, mod, "vg_grad_reduce"));
            }
            VG_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&fill_blocks_per_cu, fill, block, 0));
        }
        return CVG_OK;
    }
    ~cvg_program() {
        if (mod) (void)hipModuleUnload(mod);
    }
};

// ---------------------------------------------------------------------------
// State: device buffers of one integral.
// ---------------------------------------------------------------------------
template <typename T>
static int dalloc(T** p, long long count) {
    if (count < 1) count = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)count);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(CVG_ERR_ALLOC, std::string("hipMalloc of ") + std::to_string(sizeof(T) * count) +
                                       " bytes: " + hipGetErrorString(e));
    }
    return CVG_OK;
}

struct cvg_state {
    cvg_program* prog = nullptr;
    cvg_params p{};
    std::vector<double> domain_host;
    long long C = 0, nb = 0, s_cap = 0, run_cap = 0, win_cap = 0;
    int K = kDefaultRun, k_shift = 4;
    double V = 0.0;
    int nh0 = 2;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = -1;
    // device buffers
    double* x_edges = nullptr;
    double2* map64 = nullptr;
    float2* map32 = nullptr;
    double* weights = nullptr;
    unsigned* counts = nullptr;
    double* refine_scratch = nullptr;
    int* nh = nullptr;
    long long* nh_sum = nullptr;
    double* dh = nullptr;
    double2* JS = nullptr;  // per-cube (sum J f, sum (J f)^2)
    int* run_cube = nullptr;
    double2* win_first = nullptr;
    double2* win_last = nullptr;
    double* part = nullptr;
    long long* bsum = nullptr;
    long long* boff = nullptr;
    unsigned* sync = nullptr;  // arrival counters of the finalize [0] and block-scan [1] launches
    long long* s_total = nullptr;
    int* it_p = nullptr;
    double* dsum = nullptr;
    int* err = nullptr;
    double* hist = nullptr;
    double* domain = nullptr;
    void* theta = nullptr;
    int* fold_list = nullptr;   // cubes whose samples span more than 64 windows
    int* fold_count = nullptr;  // [0] fold list length, [1] big-run list length
    int fold_cap = 0;
    long long* big_list = nullptr;  // run-table chunks of cubes holding many runs
    int big_cap = 0;
    int* bw_fold_list = nullptr;
    int* bw_fold_count = nullptr;
    int iterations_done = 0;
    double setup_ms = 0.0, loop_ms = 0.0;
    // recorded forward iterations (for the backward replay): map edges and n_h
    // of every iteration >= skip, written by a kernel that reads the counter
    int record = 0;
    double* snap_x = nullptr;   // [max_it - skip][d*(n+1)]
    int* snap_nh = nullptr;     // [max_it - skip][C]
    // backward scratch
    long long* bw_nh_sum = nullptr;
    long long* bw_s_total = nullptr;
    double* bw_x = nullptr;
    double2* bw_map64 = nullptr;
    float2* bw_map32 = nullptr;
    double* bw_part = nullptr;  // [2 + 2 n_theta][bw_units] unit partials, then the totals
    int bw_units = 0, bw_rows_per_unit = 0;
    void* bw_theta = nullptr;   // parameters of a replay at another theta (n_theta doubles)
    unsigned char* bw_keys = nullptr;  // [2 C] size classes, then sorted
    int* bw_idx = nullptr;             // cube indices, replay order, lanes, lane starts, wave starts
    void* bw_sort_tmp = nullptr;       // radix-sort workspace
    std::vector<hipEvent_t> bw_evs;  // 2 per recordable iteration: around its replay kernels
    double bw_replay_ms = 0.0;  // device time of the last backward's replay kernels
    std::vector<char> recorded;  // per iteration: snapshot taken
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // optional per-iteration timing (eager launches): fill and whole iteration
    int timing = 0;
    std::vector<hipEvent_t> tev;  // 3 events per iteration: start, after fill, end
    std::vector<double> fill_ms, iter_ms;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;

    ~cvg_state() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (hipEvent_t e : bw_evs)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : tev) (void)hipEventDestroy(e);
        void* bufs[] = {x_edges, map64, map32, weights, counts, refine_scratch, nh, nh_sum, dh,
                        JS, run_cube, win_first, win_last, part, bsum, boff, sync, s_total,
                        it_p, dsum, err, hist, domain, theta, fold_list, fold_count, big_list, bw_fold_list,
                        bw_fold_count, snap_x, snap_nh, bw_nh_sum,
                        bw_s_total, bw_x, bw_map64, bw_map32, bw_part, bw_theta, bw_keys, bw_idx,
                        bw_sort_tmp};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }

    const void* map_for_fill() const {
        return p.precision == CVG_FP32C ? (const void*)map32 : (const void*)map64;
    }

    int launch_fill() {
        int ncu = 0;
        VG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
        const unsigned grid = (unsigned)(ncu * prog->fill_blocks_per_cu);
        const void* map = map_for_fill();
        double vol = 1.0;
        if (p.has_domain)
            for (int i = 0; i < p.n_dim; ++i) vol *= domain_host[p.n_dim + i];
        unsigned seed = p.seed, sid = p.stream_id;
        // one pass per group of prog->bin_dims dimensions (a single pass when
        // all bins fit in LDS or the bins are global); the first stores the
        // per-cube sums
        const int step = prog->lds_bins ? prog->bin_dims : p.n_dim;
        for (int lo = 0; lo < p.n_dim; lo += step) {
            int bin_lo = lo, write_sums = lo == 0 ? 1 : 0;
            void* args[] = {(void*)&map,      (void*)&nh_sum,  (void*)&run_cube, (void*)&s_total,
                            (void*)&it_p,     (void*)&JS,      (void*)&win_first,
                            (void*)&win_last, (void*)&weights, (void*)&counts,   (void*)&domain,
                            (void*)&vol,      (void*)&theta,   (void*)&seed,     (void*)&sid,
                            (void*)&bin_lo,   (void*)&write_sums};
            VG_HIP(hipModuleLaunchKernel(prog->fill, grid, 1, 1, (unsigned)prog->block, 1, 1, 0, stream,
                                         args, nullptr));
        }
        return CVG_OK;
    }

    int launch_refine() {
        VG_HIP(vg_launch_refine(stream, p.n_dim, p.n_intervals, p.alpha, x_edges, weights, counts,
                                refine_scratch, map64, map32));
        return CVG_OK;
    }

    // The results launch also refines the map (d extra blocks) when the
    // refinement's LDS arrays are small; measured on the 8-D double Gaussian,
    // the refinement (30 us on its own, latency-bound) then runs in the shadow
    // of the per-cube results.
    bool refine_in_results() const { return vg_results_refine_lds(p.n_intervals) <= 32 * 1024; }

    int launch_results(bool with_refine = false) {
        VG_HIP(vg_launch_bigfold(stream, fold_list, fold_count, fold_cap, nh_sum, run_cube, s_total,
                                 win_first, win_last, 6 + k_shift, JS));
        if (with_refine)
            VG_HIP(vg_launch_results_refine(stream, (int)C, V, p.beta, nh, JS, dh, part, p.n_dim, p.n_intervals,
                                            p.alpha, x_edges, weights, counts, refine_scratch, map64, map32));
        else
            VG_HIP(vg_launch_results(stream, (int)C, V, p.beta, nh, JS, dh, part));
        VG_HIP(vg_launch_finalize(stream, nb, part, sync, s_total, it_p, p.max_it, hist, dsum, err));
        return CVG_OK;
    }

    int launch_realloc(int mode, int advance) {
        VG_HIP(vg_launch_alloc(stream, (int)C, mode, nh0, dh, dsum, p.n_eval_it, nh, bsum));
        VG_HIP(vg_launch_scan(stream, (int)C, nh, bsum, boff, sync + 1, s_total, s_cap, it_p, advance, k_shift,
                              run_cap, nh_sum, run_cube, err, fold_list, fold_count, fold_cap, big_list,
                              big_cap));
        return CVG_OK;
    }

    int launch_realloc_copy();

    int enqueue_iteration(hipEvent_t e_start = nullptr, hipEvent_t e_fill = nullptr,
                          hipEvent_t e_end = nullptr) {
        int rc;
        if (e_start) VG_HIP(hipEventRecord(e_start, stream));
        if (record)
            VG_HIP(vg_launch_snapshot(stream, it_p, p.skip, p.max_it, (int)C, p.n_dim * (p.n_intervals + 1),
                                      x_edges, nh, snap_x, snap_nh));
        if ((rc = launch_fill())) return rc;
        if (e_fill) VG_HIP(hipEventRecord(e_fill, stream));
        if (refine_in_results()) {
            if ((rc = launch_results(true))) return rc;
        } else {
            if ((rc = launch_refine())) return rc;
            if ((rc = launch_results())) return rc;
        }
        if ((rc = launch_realloc(1, 1))) return rc;
        if (e_end) VG_HIP(hipEventRecord(e_end, stream));
        return CVG_OK;
    }

    int check_err() {
        int h_err = 0;
        VG_HIP(hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, stream));
        VG_HIP(hipStreamSynchronize(stream));
        if (h_err & VG_ERR_CAPACITY)
            return fail(CVG_ERR_CAPACITY, "sample allocation exceeded the state's buffers");
        if (h_err & VG_ERR_HIST)
            return fail(CVG_ERR_CAPACITY, "more iterations than max_it were run on this state");
        return CVG_OK;
    }
};

// Layout (nh_sum, run_cube) from the n_h already on the device.
int cvg_state::launch_realloc_copy() {
    VG_HIP(vg_launch_layout(stream, (int)C, nh, bsum, boff, sync + 1, s_total, s_cap, it_p, k_shift, run_cap, nh_sum,
                            run_cube, err, fold_list, fold_count, fold_cap, big_list, big_cap));
    return CVG_OK;
}

// ---------------------------------------------------------------------------
// Backward replay (no reference counterpart), shared by cvg_state_backward
// (buffers owned by the state) and cvg_state_backward_noalloc (buffers in the
// library's static scratch array, vg_noalloc.hip).
// ---------------------------------------------------------------------------
int vg_fail(int code, const char* msg) { return fail(code, msg); }

int vg_state_device(const cvg_state* st) { return st ? st->device : -1; }

int vg_backward_layout(const cvg_state* st, vg_bw_layout* out) {
    if (!st || !out) return fail(CVG_ERR_INVALID, "state is NULL");
    const cvg_program* prog = st->prog;
    if (!prog->has_grad || !prog->grad)
        return fail(CVG_ERR_INVALID, "program was compiled without the gradient kernel (option grad)");
    bool any = false;
    for (int k = st->p.skip; k < st->iterations_done; ++k) any = any || st->recorded[k];
    if (!any) return fail(CVG_ERR_INVALID, "no recorded iterations: enable recording before running");
    const int P = st->p.n_theta, d = st->p.n_dim, n = st->p.n_intervals;
    // waves of the replay layout (at most vg_grad_max_waves for any allocation
    // of the state's sample capacity), grouped so that at most kGradMaxUnits
    // units hold partials
    out->waves = vg_grad_max_waves(st->C, st->s_cap);
    const long long spu = (out->waves + kGradMaxUnits - 1) / kGradMaxUnits;
    out->slots_per_unit = (int)(spu > 0 ? spu : 1);
    out->units = (int)((out->waves + out->slots_per_unit - 1) / out->slots_per_unit);
    const long long slots = st->p.max_it - st->p.skip;
    out->part = ((long long)out->units + (slots > 0 ? slots : 1)) * (2LL + 2LL * P);
    out->theta = P;
    out->nh_sum = st->C + 1;
    out->x = (long long)d * (n + 1);
    out->map = (long long)d * n;
    out->cubes = st->C;
    out->sort_tmp = (long long)vg_grad_order_tmp_bytes((int)st->C);
    if (out->sort_tmp <= 0) return fail(CVG_ERR_HIP, "radix sort workspace query failed");
    return CVG_OK;
}

int vg_backward_run(cvg_state* st, const double* theta, const vg_bw_buffers& b, const vg_bw_layout& L,
                    const hipEvent_t* evs, double* grad, double* grad_var, double* value, double* value_var) {
    cvg_program* prog = st->prog;
    const int P = st->p.n_theta, d = st->p.n_dim, n = st->p.n_intervals;
    const int NV = 2 + 2 * P;
    std::vector<int> its;
    for (int k = st->p.skip; k < st->iterations_done; ++k)
        if (st->recorded[k]) its.push_back(k);
    // parameters for the replay: given theta, or the forward's
    void* th = st->theta;
    if (theta) {
        if (st->p.precision == CVG_FP32C) {
            std::vector<float> f(theta, theta + P);
            VG_HIP(hipMemcpyAsync(b.theta, f.data(), 4 * (size_t)P, hipMemcpyHostToDevice, st->stream));
            VG_HIP(hipStreamSynchronize(st->stream));
        } else {
            VG_HIP(hipMemcpyAsync(b.theta, theta, 8 * (size_t)P, hipMemcpyHostToDevice, st->stream));
        }
        th = b.theta;
    }
    std::vector<double> hist((size_t)VG_HIST * st->iterations_done);
    VG_HIP(hipMemcpyAsync(hist.data(), st->hist, sizeof(double) * hist.size(), hipMemcpyDeviceToHost, st->stream));
    const long long nx = (long long)d * (n + 1);
    std::vector<double> tot((size_t)NV * its.size());
    double W = 0.0, val = 0.0, vvar = 0.0;
    std::vector<double> g((size_t)P, 0.0), gv((size_t)P, 0.0);
    int C = (int)st->C;
    double vol = 1.0;
    if (st->p.has_domain)
        for (int i = 0; i < d; ++i) vol *= st->domain_host[d + i];
    unsigned seed = st->p.seed, sid = st->p.stream_id;
    const int nu = L.units;
    double* const totals = b.part + (long long)nu * NV;  // NV per replayed iteration
    for (size_t idx = 0; idx < its.size(); ++idx) {
        const int k = its[idx];
        const long long slot = k - st->p.skip;
        VG_HIP(hipMemcpyAsync(b.x, st->snap_x + slot * nx, sizeof(double) * nx, hipMemcpyDeviceToDevice, st->stream));
        VG_HIP(vg_launch_tables(st->stream, d, n, b.x, b.map64, b.map32));
        const int* snh = st->snap_nh + slot * st->C;
        VG_HIP(vg_launch_layout(st->stream, C, snh, st->bsum, st->boff, st->sync + 1, b.s_total, (1LL << 62),
                                st->it_p, st->k_shift, 0, b.nh_sum, nullptr, st->err, b.fold_list,
                                b.fold_count, 1, nullptr, 0));
        VG_HIP(vg_launch_grad_order(st->stream, C, snh, b.keys, b.keys_sorted, b.idx, b.order, b.lanes,
                                    b.lane_start, b.wave_first, b.sort_tmp, (size_t)L.sort_tmp));
        const void* map = st->p.precision == CVG_FP32C ? (const void*)b.map32 : (const void*)b.map64;
        unsigned itu = (unsigned)k;
        double V = st->V;
        int spu = L.slots_per_unit, nuk = nu;
        double* part = b.part;
        long long* nhs = b.nh_sum;
        int *order = b.order, *lane_start = b.lane_start, *wave_first = b.wave_first;
        void* args[] = {(void*)&map, (void*)&snh, (void*)&nhs, (void*)&order, (void*)&lane_start,
                        (void*)&wave_first, (void*)&C, (void*)&spu, (void*)&itu, (void*)&V, (void*)&st->domain,
                        (void*)&vol, (void*)&th, (void*)&seed, (void*)&sid, (void*)&part};
        if (evs) VG_HIP(hipEventRecord(evs[2 * idx], st->stream));
        VG_HIP(hipModuleLaunchKernel(prog->grad, (unsigned)nu, 1, 1, 64, 1, 1, 0, st->stream, args, nullptr));
        double* out = totals + (long long)idx * NV;
        void* rargs[] = {(void*)&part, (void*)&nuk, (void*)&out};
        VG_HIP(hipModuleLaunchKernel(prog->grad_reduce, (unsigned)NV, 1, 1, 1024, 1, 1, 0, st->stream, rargs,
                                     nullptr));
        if (evs) VG_HIP(hipEventRecord(evs[2 * idx + 1], st->stream));
    }
    VG_HIP(hipMemcpyAsync(tot.data(), totals, sizeof(double) * tot.size(), hipMemcpyDeviceToHost, st->stream));
    VG_HIP(hipStreamSynchronize(st->stream));
    double replay_ms = 0.0;
    for (size_t idx = 0; idx < its.size(); ++idx) {
        if (evs) {
            float kms = 0.0f;
            VG_HIP(hipEventElapsedTime(&kms, evs[2 * idx], evs[2 * idx + 1]));
            replay_ms += kms;
        }
        const double* t = tot.data() + idx * NV;
        const double w = 1.0 / hist[(size_t)VG_HIST * its[idx] + 1];  // forward weight, held fixed
        W += w;
        val += w * t[0];
        vvar += w * w * t[1];
        for (int q = 0; q < P; ++q) {
            g[q] += w * t[2 + q];
            gv[q] += w * w * t[2 + P + q];
        }
    }
    if (evs) st->bw_replay_ms = replay_ms;
    if (value) *value = val / W;
    if (value_var) *value_var = vvar / (W * W);
    for (int q = 0; q < P; ++q) {
        if (grad) grad[q] = g[q] / W;
        if (grad_var) grad_var[q] = gv[q] / (W * W);
    }
    return st->check_err();
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* cvg_last_error(void) { return g_err.c_str(); }

int cvg_device_info(int* n_cu, int* clock_khz, char* arch, int arch_len) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(CVG_ERR_NODEVICE, "no HIP device visible");
    int dev;
    VG_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    VG_HIP(hipGetDeviceProperties(&prop, dev));
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (clock_khz) *clock_khz = prop.clockRate;
    if (arch && arch_len > 0) {
        std::snprintf(arch, (size_t)arch_len, "%s", prop.gcnArchName);
    }
    return CVG_OK;
}

int cvg_program_create(const char* integrand_src, const cvg_params* shape, const char* options,
                       cvg_program** out) {
    if (!out) return fail(CVG_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!integrand_src) return fail(CVG_ERR_INVALID, "integrand source is NULL");
    int rc = validate(shape);
    if (rc) return rc;
    auto prog = std::make_unique<cvg_program>();
    prog->d = shape->n_dim;
    prog->nbins = shape->n_intervals;
    prog->nstrat = shape->n_strat;
    prog->precision = shape->precision;
    prog->K = shape->run_len > 0 ? shape->run_len : kDefaultRun;
    // Sampling-loop schedule, measured (fill kernel, MI355X):
    //  * d <= 8: one sample per lane step with the next sample's map entries
    //    requested before the current one is evaluated, FP32 map read through
    //    the caches, so the LDS serves only the bin atomics that bound the
    //    FP32 kernel.  8-D double Gaussian, 1e8 samples: FP32 1.29 ms against
    //    1.35 ms with the map in LDS and no prefetch; FP64 1.66 ms against
    //    1.70 ms with two samples per step.
    //  * d > 8: the prefetched sample's state does not fit next to the
    //    current one (16-D oscillatory, FP64: 85 VGPRs spilled, fill 8.8 ms
    //    against 4.8 ms without the prefetch; FP32 3.42 against 3.33 ms), so
    //    one sample per step.
    //  * The FP32 map is read through the caches in every shape: staged in
    //    LDS beside the bins it adds d LDS reads per sample to an LDS that the
    //    bin atomics already keep busy (16-D oscillatory, 512 bins, 2e8
    //    samples: fill 3.66 ms with the 64 KB map in LDS, 3.20 ms through the
    //    caches; option lds_map=1 stages it).
    // 16 waves per CU for the FP32 form, 12 for the FP64 form (168 VGPRs).
    prog->block = prog->precision == CVG_FP32C ? kFillBlock : 768;
    prog->pair = 0;
    prog->pipe = prog->d <= 8 ? 1 : 0;
    prog->has_domain = shape->has_domain ? 1 : 0;
    prog->n_theta = shape->n_theta;
    const long long dn = (long long)prog->d * prog->nbins;
    long long bin_bytes = dn * 12;
    const long long map_bytes = prog->precision == CVG_FP32C ? dn * 8 : 0;
    // Bins of as many dimensions as fit in LDS per sampling pass; when not all
    // fit, the fill runs once per group of dimensions (the same samples drawn
    // again) instead of adding every sample to global bins (16-D, 1024 bins,
    // 2e8 samples: 170 ms per iteration with global atomics).
    auto set_bin_dims = [&](long long bytes_per_bin) {
        const long long fit = kLdsBytes / (bytes_per_bin * prog->nbins);
        prog->bin_dims = (int)(fit < prog->d ? fit : prog->d);
        prog->lds_bins = prog->bin_dims >= 1;
        if (prog->bin_dims < 1) prog->bin_dims = 1;
    };
    set_bin_dims(12);
    prog->lds_map = 0;

    // Launch-shape knobs in the options string (consumed here, not passed to
    // hipRTC): lds_bins=0|1, lds_map=0|1, bin32=0|1, block=64..1024 (multiple
    // of 64), pair=0|1 (two samples per lane step in the sampling loop),
    // pipe=0|1 (one sample per step; 1: the next sample's map entries are
    // requested before the current sample is evaluated), grad=0|1.  All are
    // read first and applied in a fixed order (bin32, then lds_bins, then
    // lds_map), so the result does not depend on the order of the tokens.
    std::string rtc_extra;
    int opt_lds_bins = -1, opt_lds_map = -1, opt_bin32 = -1, opt_copies = -1, opt_copy_dim0 = -1;
    if (options) {
        std::istringstream ss(options);
        std::string tok;
        while (ss >> tok) {
            const size_t eq = tok.find('=');
            const std::string key = eq == std::string::npos ? tok : tok.substr(0, eq);
            const int val = eq == std::string::npos ? 1 : std::atoi(tok.c_str() + eq + 1);
            if (key == "lds_bins") {
                opt_lds_bins = val ? 1 : 0;
            } else if (key == "lds_map") {
                opt_lds_map = val ? 1 : 0;
            } else if (key == "bin32") {
                if (val && prog->precision != CVG_FP32C) return fail(CVG_ERR_INVALID, "bin32 is for the FP32 form");
                opt_bin32 = val ? 1 : 0;
            } else if (key == "bin_copies") {
                if (val < 1 || val > 8) return fail(CVG_ERR_INVALID, "bin_copies must be in 1..8");
                opt_copies = val;
            } else if (key == "copy_dim0") {
                if (val < 0 || val >= prog->d) return fail(CVG_ERR_INVALID, "copy_dim0 must be in 0..n_dim-1");
                opt_copy_dim0 = val;
            } else if (key == "grad") {
                if (shape->n_theta < 1) return fail(CVG_ERR_INVALID, "grad needs n_theta >= 1");
                prog->has_grad = val ? 1 : 0;
            } else if (key == "block") {
                if (val < 64 || val > 1024 || val % 64) return fail(CVG_ERR_INVALID, "block must be 64..1024, multiple of 64");
                prog->block = val;
            } else if (key == "pair") {
                prog->pair = val ? 1 : 0;
                if (prog->pair) prog->pipe = 0;
            } else if (key == "pipe") {
                if (val < 0 || val > 1) return fail(CVG_ERR_INVALID, "pipe must be 0 or 1");
                prog->pipe = val;
                if (prog->pipe) prog->pair = 0;
            } else {
                rtc_extra += tok + " ";
            }
        }
    }
    if (opt_bin32 >= 0) {
        prog->bin32 = opt_bin32;
        bin_bytes = dn * (prog->bin32 ? 8 : 12);
        set_bin_dims(prog->bin32 ? 8 : 12);
        if (!prog->lds_bins || bin_bytes + map_bytes > kLdsBytes) prog->lds_map = 0;
    }
    if (opt_lds_bins >= 0) {
        if (opt_lds_bins && (long long)prog->nbins * (prog->bin32 ? 8 : 12) > kLdsBytes)
            return fail(CVG_ERR_INVALID, "the bins of one dimension do not fit in LDS");
        prog->lds_bins = opt_lds_bins;
        if (!prog->lds_bins) prog->lds_map = 0;
    }
    if (opt_lds_map >= 0) {
        if (opt_lds_map && (prog->precision != CVG_FP32C || !prog->lds_bins || bin_bytes + map_bytes > kLdsBytes))
            return fail(CVG_ERR_INVALID, "map does not fit in LDS next to the bins");
        prog->lds_map = opt_lds_map;
    }
    // Copies of the LDS bins in the dimensions whose stratum digit a window's
    // cubes share (from the first dimension i with n_strat^i >= 256: a window
    // of 64 K samples holds fewer cubes than that digit's period once the
    // allocation has adapted) when a stratum spans few bins: the lanes of a
    // wave then add into the bins of one stratum, and lane l adds into copy
    // l mod copies.  As many copies (8, 4 or 2) as fit in LDS.  4-D Gaussian,
    // 1024 bins, 70 strata (14.6 bins per stratum): dimensions 2 and 3, 4
    // copies, both forms.  Up to 256 bins per stratum in the FP32 form, which
    // the LDS co-bounds: 8-D double Gaussian, 1024 bins, 8 strata: dimensions
    // 3-7, 2 copies, fill 0.4-1.5% faster in paired same-box runs; 16-D
    // oscillatory, 512 bins, 2 strata: dimensions 8-15, 2 copies, 1.9% faster.
    // The FP64 form, held by its VALU work, ran 2% slower with them at the
    // 8-D shape, so it takes copies only up to 32 bins per stratum.
    auto copies_lds = [&](int copies, int dim0) {
        return (long long)(copies - 1) * ((long long)(prog->d - dim0) * prog->nbins + 17) * (prog->bin32 ? 8 : 12);
    };
    const long long lds_used = bin_bytes + (prog->lds_map ? map_bytes : 0);
    const bool copies_ok = prog->lds_bins && prog->bin_dims == prog->d;
    prog->bin_copies = 1;
    prog->copy_dim0 = prog->d;
    const long long span_max = prog->precision == CVG_FP32C ? 256 : 32;
    if (copies_ok && prog->nbins <= span_max * prog->nstrat) {
        long long period = 1;
        int i0 = 0;
        while (i0 < prog->d && period < 256) { period *= prog->nstrat; ++i0; }
        for (int cp : {8, 4, 2}) {
            if (i0 < prog->d && lds_used + copies_lds(cp, i0) <= kLdsBytes) {
                prog->bin_copies = cp;
                prog->copy_dim0 = i0;
                break;
            }
        }
    }
    if (opt_copies >= 0 || opt_copy_dim0 >= 0) {
        const int cp = opt_copies >= 0 ? opt_copies : (prog->bin_copies > 1 ? prog->bin_copies : 1);
        const int i0 = opt_copy_dim0 >= 0 ? opt_copy_dim0 : (prog->copy_dim0 < prog->d ? prog->copy_dim0 : prog->d - 1);
        if (cp > 1 && (!copies_ok || lds_used + copies_lds(cp, i0) > kLdsBytes))
            return fail(CVG_ERR_INVALID, "bin copies need all bins in LDS and room for the copies");
        prog->bin_copies = cp;
        prog->copy_dim0 = cp > 1 ? i0 : prog->d;
    }
    options = rtc_extra.c_str();
    const double t0 = now_ms();
    hiprtcProgram rtc;
    const char* hdr_src[3] = {kVgPhiloxSrc, kVgMathSrc, integrand_src};
    const char* hdr_name[3] = {"vg_philox.h", "vg_math.h", "vg_user.h"};
    if (hiprtcCreateProgram(&rtc, kVgFillSrc, "vg_fill.hip", 3, hdr_src, hdr_name) != HIPRTC_SUCCESS)
        return fail(CVG_ERR_COMPILE, "hiprtcCreateProgram failed");
    std::vector<std::string> opts = {
        "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-DVG_NDIM=" + std::to_string(prog->d), "-DVG_NBINS=" + std::to_string(prog->nbins),
        "-DVG_NSTRAT=" + std::to_string(prog->nstrat) + "u", "-DVG_K=" + std::to_string(prog->K),
        "-DVG_BLOCK=" + std::to_string(prog->block),
        "-DVG_FP32=" + std::to_string(prog->precision == CVG_FP32C ? 1 : 0),
        "-DVG_LDS_BINS=" + std::to_string(prog->lds_bins),
        "-DVG_BIN_DIMS=" + std::to_string(prog->bin_dims),
        "-DVG_BIN_COPIES=" + std::to_string(prog->bin_copies),
        "-DVG_COPY_DIM0=" + std::to_string(prog->copy_dim0),
        "-DVG_LDS_MAP=" + std::to_string(prog->lds_map),
        "-DVG_HAS_DOMAIN=" + std::to_string(prog->has_domain),
        "-DVG_GRAD=" + std::to_string(prog->has_grad),
        "-DVG_BIN32=" + std::to_string(prog->bin32),
        "-DVG_PAIR=" + std::to_string(prog->pair),
        "-DVG_PIPE=" + std::to_string(prog->pipe),
        "-DVG_NTHETA=" + std::to_string(prog->n_theta)};
    if (options) {
        std::istringstream ss(options);
        std::string tok;
        while (ss >> tok) opts.push_back(tok);
    }

    std::vector<const char*> copts;
    for (auto& o : opts) copts.push_back(o.c_str());
    hiprtcResult cr = hiprtcCompileProgram(rtc, (int)copts.size(), copts.data());
    size_t log_size = 0;
    hiprtcGetProgramLogSize(rtc, &log_size);
    if (log_size > 1) {
        prog->log.resize(log_size);
        hiprtcGetProgramLog(rtc, &prog->log[0]);
        while (!prog->log.empty() && prog->log.back() == '\0') prog->log.pop_back();
    }
    if (cr != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&rtc);
        *out = prog.release();
        return fail(CVG_ERR_COMPILE, std::string("hipRTC compile failed: ") + hiprtcGetErrorString(cr) +
                                         "\n" + (*out)->log);
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(rtc, &code_size);
    prog->code.resize(code_size);
    hiprtcGetCode(rtc, prog->code.data());
    hiprtcDestroyProgram(&rtc);
    prog->compile_ms = now_ms() - t0;
    *out = prog.release();
    return CVG_OK;
}

const char* cvg_program_log(const cvg_program* prog) { return prog ? prog->log.c_str() : ""; }

long long cvg_program_code_size(const cvg_program* prog) {
    return prog ? (long long)prog->code.size() : 0;
}

long long cvg_program_get_code(const cvg_program* prog, void* buf, long long cap) {
    if (!prog || !buf || cap <= 0) return 0;
    const long long n = (long long)prog->code.size() < cap ? (long long)prog->code.size() : cap;
    std::memcpy(buf, prog->code.data(), (size_t)n);
    return n;
}

void cvg_program_destroy(cvg_program* prog) { delete prog; }

int cvg_state_create(cvg_program* prog, const cvg_params* params, cvg_state** out) {
    if (!out) return fail(CVG_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!prog || prog->code.empty()) return fail(CVG_ERR_INVALID, "program is NULL or not compiled");
    int rc = validate(params);
    if (rc) return rc;
    const int K = params->run_len > 0 ? params->run_len : kDefaultRun;
    if (params->n_dim != prog->d || params->n_intervals != prog->nbins || params->n_strat != prog->nstrat ||
        params->precision != prog->precision || K != prog->K || params->n_theta != prog->n_theta ||
        (params->has_domain ? 1 : 0) != prog->has_domain)
        return fail(CVG_ERR_INVALID, "params do not match the program's compiled shape");
    if (params->max_it < 1) return fail(CVG_ERR_INVALID, "max_it must be >= 1");
    if (params->n_eval_it < 1) return fail(CVG_ERR_INVALID, "n_eval_it must be >= 1");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(CVG_ERR_NODEVICE, "no HIP device visible");
    if ((rc = prog->load())) return rc;

    const double t0 = now_ms();
    auto st = std::make_unique<cvg_state>();
    st->prog = prog;
    st->p = *params;
    st->K = K;
    st->k_shift = 0;
    while ((1 << st->k_shift) < K) ++st->k_shift;
    VG_HIP(hipGetDevice(&st->device));
    if (params->has_domain) {
        st->domain_host.assign(params->domain, params->domain + 2 * params->n_dim);
        st->p.domain = nullptr;
    }
    const int d = params->n_dim, n = params->n_intervals;
    long long C = 0;
    checked_cubes(params->n_strat, d, &C);
    st->C = C;
    st->nb = (C + 1023) / 1024;
    st->V = std::pow(1.0 / params->n_strat, d);
    // Reference: int nh_s = 1.0 / (double) n_cubes * n_eval_it; max(2, .)  (cuda_wrapper.cu:475-480)
    {
        const double v = 1.0 / (double)C * (double)params->n_eval_it;
        int nh0 = v != v ? 0 : (v >= 2147483647.0 ? 2147483647 : (int)v);
        st->nh0 = nh0 < 2 ? 2 : nh0;
    }
    // Sum of n_h <= N + 2C; a margin covers rounding of d_h / D.
    st->s_cap = params->n_eval_it + 2 * C + 64LL * K + 1024;
    if ((long long)st->nh0 * C > st->s_cap) st->s_cap = (long long)st->nh0 * C;
    st->run_cap = (st->s_cap + K - 1) / K + 1;
    st->win_cap = (st->s_cap + 64LL * K - 1) / (64LL * K) + 1;

    if ((rc = dalloc(&st->x_edges, (long long)d * (n + 1)))) return rc;
    if ((rc = dalloc(&st->map64, (long long)d * n))) return rc;
    if ((rc = dalloc(&st->map32, (long long)d * n))) return rc;
    if ((rc = dalloc(&st->weights, (long long)d * n))) return rc;
    if ((rc = dalloc(&st->counts, (long long)d * n))) return rc;
    if ((rc = dalloc(&st->refine_scratch, vg_refine_scratch_doubles(d, n)))) return rc;
    if ((rc = dalloc(&st->nh, C))) return rc;
    if ((rc = dalloc(&st->nh_sum, C + 1))) return rc;
    if ((rc = dalloc(&st->dh, C))) return rc;
    if ((rc = dalloc(&st->JS, C))) return rc;
    if ((rc = dalloc(&st->run_cube, st->run_cap))) return rc;
    if ((rc = dalloc(&st->win_first, st->win_cap))) return rc;
    if ((rc = dalloc(&st->win_last, st->win_cap))) return rc;
    // block partials, then one per 1024 blocks (the group level of the trees / scans)
    const long long ng = (st->nb + 1023) / 1024;
    if ((rc = dalloc(&st->part, (long long)VG_NRES * (st->nb + ng)))) return rc;
    if ((rc = dalloc(&st->bsum, st->nb + ng))) return rc;
    if ((rc = dalloc(&st->boff, st->nb + ng))) return rc;
    if ((rc = dalloc(&st->sync, 2))) return rc;
    if ((rc = dalloc(&st->s_total, 1))) return rc;
    if ((rc = dalloc(&st->it_p, 1))) return rc;
    if ((rc = dalloc(&st->dsum, 1))) return rc;
    if ((rc = dalloc(&st->err, 1))) return rc;
    if ((rc = dalloc(&st->hist, (long long)VG_HIST * params->max_it))) return rc;
    if ((rc = dalloc(&st->domain, 2LL * d))) return rc;
    st->fold_cap = (int)(st->s_cap / (64LL * 64 * K) + 2);
    if ((rc = dalloc(&st->fold_list, st->fold_cap))) return rc;
    if ((rc = dalloc(&st->fold_count, 2))) return rc;
    st->big_cap = (int)vg_big_list_cap(st->run_cap);
    if ((rc = dalloc(&st->big_list, 2LL * st->big_cap))) return rc;
    if ((rc = dalloc((char**)&st->theta, 8LL * (params->n_theta > 0 ? params->n_theta : 1)))) return rc;

    VG_HIP(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
    st->own_stream = true;
    VG_HIP(hipEventCreate(&st->ev0));
    VG_HIP(hipEventCreate(&st->ev1));
    hipStream_t s = st->stream;
    if (params->has_domain)
        VG_HIP(hipMemcpyAsync(st->domain, st->domain_host.data(), sizeof(double) * 2 * d,
                              hipMemcpyHostToDevice, s));
    VG_HIP(hipMemsetAsync(st->it_p, 0, sizeof(int), s));
    VG_HIP(hipMemsetAsync(st->err, 0, sizeof(int), s));
    VG_HIP(hipMemsetAsync(st->sync, 0, 2 * sizeof(unsigned), s));
    VG_HIP(hipMemsetAsync(st->theta, 0, 8LL * (params->n_theta > 0 ? params->n_theta : 1), s));
    VG_HIP(vg_launch_init_map(s, d, n, st->x_edges, st->map64, st->map32, st->weights, st->counts));
    if ((rc = st->launch_realloc(0, 0))) return rc;
    VG_HIP(hipStreamSynchronize(s));
    st->recorded.assign((size_t)params->max_it, 0);
    st->setup_ms = now_ms() - t0;
    *out = st.release();
    return CVG_OK;
}

void cvg_state_destroy(cvg_state* st) {
    if (st && st->stream) (void)hipStreamSynchronize(st->stream);
    delete st;
}

long long cvg_state_n_cubes(const cvg_state* st) { return st ? st->C : 0; }

int cvg_state_set_stream(cvg_state* st, void* hip_stream) {
    if (!st) return fail(CVG_ERR_INVALID, "state is NULL");
    VG_HIP(hipStreamSynchronize(st->stream));
    if (st->graph_exec) { (void)hipGraphExecDestroy(st->graph_exec); st->graph_exec = nullptr; }
    if (st->graph) { (void)hipGraphDestroy(st->graph); st->graph = nullptr; }
    if (hip_stream) {
        if (st->own_stream) VG_HIP(hipStreamDestroy(st->stream));
        st->stream = (hipStream_t)hip_stream;
        st->own_stream = false;
    } else if (!st->own_stream) {
        VG_HIP(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
        st->own_stream = true;
    }
    return CVG_OK;
}

int cvg_state_set_theta(cvg_state* st, const double* theta, int n_theta) {
    if (!st) return fail(CVG_ERR_INVALID, "state is NULL");
    if (n_theta != st->p.n_theta || (n_theta > 0 && !theta))
        return fail(CVG_ERR_INVALID, "theta size does not match n_theta");
    if (n_theta == 0) return CVG_OK;
    if (st->p.precision == CVG_FP32C) {
        std::vector<float> f(theta, theta + n_theta);
        VG_HIP(hipMemcpyAsync(st->theta, f.data(), sizeof(float) * n_theta, hipMemcpyHostToDevice, st->stream));
        VG_HIP(hipStreamSynchronize(st->stream));
    } else {
        VG_HIP(hipMemcpyAsync(st->theta, theta, sizeof(double) * n_theta, hipMemcpyHostToDevice, st->stream));
        VG_HIP(hipStreamSynchronize(st->stream));
    }
    return CVG_OK;
}

int cvg_state_run(cvg_state* st, int n_iter, int use_graph) {
    if (!st) return fail(CVG_ERR_INVALID, "state is NULL");
    if (n_iter < 0) return fail(CVG_ERR_INVALID, "n_iter must be >= 0");
    if (st->iterations_done + n_iter > st->p.max_it)
        return fail(CVG_ERR_CAPACITY, "more iterations than max_it requested");
    if (n_iter == 0) return CVG_OK;
    int rc;
    if (use_graph && !st->graph_exec) {
        VG_HIP(hipStreamBeginCapture(st->stream, hipStreamCaptureModeThreadLocal));
        rc = st->enqueue_iteration();
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(st->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ce != hipSuccess) return fail(CVG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        st->graph = g;
        VG_HIP(hipGraphInstantiate(&st->graph_exec, st->graph, nullptr, nullptr, 0));
    }
    const bool timed = st->timing && !use_graph;
    if (timed) {
        while ((int)st->tev.size() < 3 * n_iter) {
            hipEvent_t e;
            VG_HIP(hipEventCreate(&e));
            st->tev.push_back(e);
        }
    }
    VG_HIP(hipEventRecord(st->ev0, st->stream));
    for (int i = 0; i < n_iter; ++i) {
        if (use_graph) {
            VG_HIP(hipGraphLaunch(st->graph_exec, st->stream));
        } else if (timed) {
            if ((rc = st->enqueue_iteration(st->tev[3 * i], st->tev[3 * i + 1], st->tev[3 * i + 2])))
                return rc;
        } else if ((rc = st->enqueue_iteration())) {
            return rc;
        }
    }
    VG_HIP(hipEventRecord(st->ev1, st->stream));
    VG_HIP(hipEventSynchronize(st->ev1));
    float ms = 0.0f;
    VG_HIP(hipEventElapsedTime(&ms, st->ev0, st->ev1));
    st->loop_ms += ms;
    for (int i = 0; i < n_iter; ++i) {
        const int k = st->iterations_done + i;
        if (k < (int)st->recorded.size()) st->recorded[k] = (char)(st->record && k >= st->p.skip);
    }
    st->iterations_done += n_iter;
    if (timed) {
        for (int i = 0; i < n_iter; ++i) {
            float a = 0.0f, b = 0.0f;
            VG_HIP(hipEventElapsedTime(&a, st->tev[3 * i], st->tev[3 * i + 1]));
            VG_HIP(hipEventElapsedTime(&b, st->tev[3 * i], st->tev[3 * i + 2]));
            st->fill_ms.push_back(a);
            st->iter_ms.push_back(b);
        }
    }
    return st->check_err();
}

int cvg_state_set_timing(cvg_state* st, int enable) {
    if (!st) return fail(CVG_ERR_INVALID, "state is NULL");
    st->timing = enable ? 1 : 0;
    return CVG_OK;
}

int cvg_state_get_timing(cvg_state* st, double* fill_ms, double* iter_ms, int cap) {
    if (!st) return fail(CVG_ERR_INVALID, "state is NULL");
    const int n = (int)st->fill_ms.size();
    for (int i = 0; i < n && i < cap; ++i) {
        if (fill_ms) fill_ms[i] = st->fill_ms[i];
        if (iter_ms) iter_ms[i] = st->iter_ms[i];
    }
    return n;
}

// The iteration record on the host: VG_HIST doubles per iteration.
static int read_history(cvg_state* st, std::vector<double>& h) {
    const int n = st->iterations_done;
    h.assign((size_t)VG_HIST * (n > 0 ? n : 1), 0.0);
    if (n > 0) {
        VG_HIP(hipMemcpyAsync(h.data(), st->hist, sizeof(double) * VG_HIST * n, hipMemcpyDeviceToHost,
                              st->stream));
        VG_HIP(hipStreamSynchronize(st->stream));
    }
    return CVG_OK;
}

int cvg_state_history(cvg_state* st, double* I, double* sigma2, long long* nevals, int cap) {
    if (!st) return fail(CVG_ERR_INVALID, "state is NULL");
    const int n = st->iterations_done;
    std::vector<double> h;
    int rc = read_history(st, h);
    if (rc) return rc;
    for (int i = 0; i < n && i < cap; ++i) {
        if (I) I[i] = h[VG_HIST * i + 0];
        if (sigma2) sigma2[i] = h[VG_HIST * i + 1];
        if (nevals) nevals[i] = (long long)h[VG_HIST * i + 3];
    }
    return n;
}

int cvg_state_history_unbiased(cvg_state* st, double* sigma2_unbiased, int cap) {
    if (!st || !sigma2_unbiased) return fail(CVG_ERR_INVALID, "NULL argument");
    const int n = st->iterations_done;
    std::vector<double> h;
    int rc = read_history(st, h);
    if (rc) return rc;
    for (int i = 0; i < n && i < cap; ++i) sigma2_unbiased[i] = h[VG_HIST * i + 4];
    return n;
}

// Inverse-variance weighted mean of iterations [skip, n) of a record of
// VG_HIST doubles per iteration (reference combine, cuda_wrapper.cu:758-766
// and :869), its error and chi2/dof; with the unbiased per-iteration
// variances, the error of that same weighted mean and chi2/dof (no reference
// counterpart).
struct vg_combined {
    double result = 0.0, error = 0.0, chi2_dof = 0.0, error_unb = 0.0, chi2_dof_unb = 0.0;
    long long tot = 0;
};
static vg_combined combine_record(const double* h, int n, int skip) {
    vg_combined c;
    for (int i = 0; i < n; ++i) c.tot += (long long)h[VG_HIST * i + 3];
    // Iterations of zero variance (an integrand the map and strata make
    // constant, e.g. f = 1) are exact: their weight 1 / sigma_k^2 is infinite,
    // and the reference's sums give inf / inf = NaN.  Here the result is the
    // mean of those iterations, with error and chi2/dof 0.
    double zsum = 0.0;
    int nz = 0;
    for (int i = skip; i < n; ++i)
        if (h[VG_HIST * i + 1] == 0.0) { zsum += h[VG_HIST * i]; ++nz; }
    if (nz > 0) {
        c.result = zsum / nz;
        return c;
    }
    double res = 0.0, sig = 0.0;
    int Kc = 0;
    for (int i = skip; i < n; ++i) {
        res += h[VG_HIST * i] / h[VG_HIST * i + 1];
        sig += 1.0 / h[VG_HIST * i + 1];
        ++Kc;
    }
    if (Kc > 0) res /= sig;
    double chi2 = 0.0, chi2u = 0.0, varu = 0.0;
    for (int i = skip; i < n; ++i) {
        const double dd = h[VG_HIST * i] - res, s2 = h[VG_HIST * i + 1], s2u = h[VG_HIST * i + 4];
        chi2 += dd * dd / s2;
        chi2u += dd * dd / s2u;
        varu += s2u / (s2 * s2);  // w_k^2 sigma_unb,k^2 with w_k = 1 / sigma_k^2
    }
    c.result = res;
    c.error = 1.0 / std::sqrt(sig);
    c.chi2_dof = Kc > 1 ? chi2 / (Kc - 1) : 0.0;
    c.error_unb = Kc > 0 ? std::sqrt(varu) / sig : 0.0;
    c.chi2_dof_unb = Kc > 1 ? chi2u / (Kc - 1) : 0.0;
    return c;
}

int cvg_state_result(cvg_state* st, cvg_result* out) {
    if (!st || !out) return fail(CVG_ERR_INVALID, "state or out is NULL");
    const int n = st->iterations_done;
    std::vector<double> h;
    int rc = read_history(st, h);
    if (rc) return rc;
    const vg_combined c = combine_record(h.data(), n, st->p.skip);
    out->result = c.result;
    out->error = c.error;
    out->chi2_dof = c.chi2_dof;
    out->error_unbiased = c.error_unb;
    out->chi2_dof_unbiased = c.chi2_dof_unb;
    out->tot_nevals = c.tot;
    out->loop_ms = st->loop_ms;
    out->compile_ms = st->prog->compile_ms;
    out->setup_ms = st->setup_ms;
    out->iterations = n;
    out->n_cubes = (int)st->C;
    out->n_strat = st->p.n_strat;
    out->n_intervals = st->p.n_intervals;
    return CVG_OK;
}

static int copy_d2h(cvg_state* st, void* dst, const void* src, size_t bytes) {
    VG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st->stream));
    VG_HIP(hipStreamSynchronize(st->stream));
    return CVG_OK;
}
static int copy_h2d(cvg_state* st, void* dst, const void* src, size_t bytes) {
    VG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st->stream));
    VG_HIP(hipStreamSynchronize(st->stream));
    return CVG_OK;
}

int cvg_state_get_map(cvg_state* st, double* x_edges) {
    if (!st || !x_edges) return fail(CVG_ERR_INVALID, "NULL argument");
    return copy_d2h(st, x_edges, st->x_edges, sizeof(double) * st->p.n_dim * (st->p.n_intervals + 1));
}

int cvg_state_set_map(cvg_state* st, const double* x_edges) {
    if (!st || !x_edges) return fail(CVG_ERR_INVALID, "NULL argument");
    const int d = st->p.n_dim, n = st->p.n_intervals;
    for (int i = 0; i < d; ++i) {
        const double* x = x_edges + (size_t)i * (n + 1);
        for (int k = 0; k < n; ++k)
            if (!(x[k + 1] > x[k])) return fail(CVG_ERR_INVALID, "map edges must be strictly increasing");
    }
    int rc = copy_h2d(st, st->x_edges, x_edges, sizeof(double) * d * (n + 1));
    if (rc) return rc;
    VG_HIP(vg_launch_tables(st->stream, d, n, st->x_edges, st->map64, st->map32));
    VG_HIP(hipStreamSynchronize(st->stream));
    return CVG_OK;
}

int cvg_state_get_nh(cvg_state* st, int* nh) {
    if (!st || !nh) return fail(CVG_ERR_INVALID, "NULL argument");
    return copy_d2h(st, nh, st->nh, sizeof(int) * st->C);
}

int cvg_state_set_nh(cvg_state* st, const int* nh) {
    if (!st || !nh) return fail(CVG_ERR_INVALID, "NULL argument");
    long long tot = 0;
    for (long long h = 0; h < st->C; ++h) {
        // the allocation rule never gives fewer than 2 (cuda_wrapper.cu:130,477),
        // and the sampling loop takes two samples per step on that guarantee
        if (nh[h] < 2) return fail(CVG_ERR_INVALID, "n_h must be >= 2");
        tot += nh[h];
    }
    if (tot > st->s_cap) return fail(CVG_ERR_CAPACITY, "sum of n_h exceeds the state's sample buffers");
    int rc = copy_h2d(st, st->nh, nh, sizeof(int) * st->C);
    if (rc) return rc;
    if ((rc = st->launch_realloc_copy())) return rc;
    VG_HIP(hipStreamSynchronize(st->stream));
    return st->check_err();
}

int cvg_state_get_dh(cvg_state* st, double* dh) {
    if (!st || !dh) return fail(CVG_ERR_INVALID, "NULL argument");
    return copy_d2h(st, dh, st->dh, sizeof(double) * st->C);
}

int cvg_state_get_cube_sums(cvg_state* st, double* jf, double* jf2) {
    if (!st || !jf || !jf2) return fail(CVG_ERR_INVALID, "NULL argument");
    std::vector<double2> js((size_t)st->C);
    int rc = copy_d2h(st, js.data(), st->JS, sizeof(double2) * st->C);
    if (rc) return rc;
    for (long long h = 0; h < st->C; ++h) { jf[h] = js[h].x; jf2[h] = js[h].y; }
    return CVG_OK;
}

int cvg_state_get_bins(cvg_state* st, double* weights, unsigned int* counts) {
    if (!st || !weights || !counts) return fail(CVG_ERR_INVALID, "NULL argument");
    const size_t dn = (size_t)st->p.n_dim * st->p.n_intervals;
    int rc = copy_d2h(st, weights, st->weights, sizeof(double) * dn);
    if (rc) return rc;
    return copy_d2h(st, counts, st->counts, sizeof(unsigned) * dn);
}

int cvg_state_set_bins(cvg_state* st, const double* weights, const unsigned int* counts) {
    if (!st || !weights || !counts) return fail(CVG_ERR_INVALID, "NULL argument");
    const size_t dn = (size_t)st->p.n_dim * st->p.n_intervals;
    int rc = copy_h2d(st, st->weights, weights, sizeof(double) * dn);
    if (rc) return rc;
    return copy_h2d(st, st->counts, counts, sizeof(unsigned) * dn);
}

int cvg_state_set_history(cvg_state* st, const double* I, const double* sigma2, const long long* nevals,
                          const double* sigma2_unbiased, int count) {
    if (!st || count < 0) return fail(CVG_ERR_INVALID, "bad argument");
    if (count > 0 && (!I || !sigma2 || !nevals)) return fail(CVG_ERR_INVALID, "NULL argument");
    if (count > st->p.max_it) return fail(CVG_ERR_INVALID, "count exceeds the state's max_it");
    if (count > 0) {
        std::vector<double> h((size_t)VG_HIST * count);
        for (int i = 0; i < count; ++i) {
            h[VG_HIST * i + 0] = I[i];
            h[VG_HIST * i + 1] = sigma2[i];
            h[VG_HIST * i + 2] = 0.0;  // D of a restored iteration is not kept
            h[VG_HIST * i + 3] = (double)nevals[i];
            h[VG_HIST * i + 4] = sigma2_unbiased ? sigma2_unbiased[i] : sigma2[i];
        }
        int rc = copy_h2d(st, st->hist, h.data(), sizeof(double) * h.size());
        if (rc) return rc;
    }
    st->iterations_done = count;
    return copy_h2d(st, st->it_p, &count, sizeof(int));
}

int cvg_state_set_iteration(cvg_state* st, int it) {
    if (!st || it < 0) return fail(CVG_ERR_INVALID, "bad argument");
    return copy_h2d(st, st->it_p, &it, sizeof(int));
}

int cvg_state_fill(cvg_state* st) {
    if (!st) return fail(CVG_ERR_INVALID, "state is NULL");
    int rc = st->launch_fill();
    if (rc) return rc;
    VG_HIP(hipStreamSynchronize(st->stream));
    return CVG_OK;
}

int cvg_state_refine(cvg_state* st) {
    if (!st) return fail(CVG_ERR_INVALID, "state is NULL");
    int rc = st->launch_refine();
    if (rc) return rc;
    VG_HIP(hipStreamSynchronize(st->stream));
    return CVG_OK;
}

int cvg_state_reallocate(cvg_state* st, const double* dh, double D) {
    if (!st || !dh) return fail(CVG_ERR_INVALID, "NULL argument");
    int rc = copy_h2d(st, st->dh, dh, sizeof(double) * st->C);
    if (rc) return rc;
    if ((rc = copy_h2d(st, st->dsum, &D, sizeof(double)))) return rc;
    if ((rc = st->launch_realloc(1, 0))) return rc;
    VG_HIP(hipStreamSynchronize(st->stream));
    return st->check_err();
}

int cvg_state_dump(cvg_state* st, int it, long long n, long long* cube, int* ids, double* x, double* jf) {
    if (!st || !cube || !ids || !x || !jf || n < 0 || it < 0) return fail(CVG_ERR_INVALID, "bad argument");
    if (n == 0) return CVG_OK;
    const int d = st->p.n_dim;
    long long* d_cube = nullptr;
    int* d_ids = nullptr;
    double *d_x = nullptr, *d_jf = nullptr;
    int rc;
    if ((rc = dalloc(&d_cube, n)) || (rc = dalloc(&d_ids, n * d)) || (rc = dalloc(&d_x, n * d)) ||
        (rc = dalloc(&d_jf, n))) {
        (void)hipFree(d_cube); (void)hipFree(d_ids); (void)hipFree(d_x); (void)hipFree(d_jf);
        return rc;
    }
    const void* map = st->map_for_fill();
    int C = (int)st->C;
    unsigned itu = (unsigned)it, seed = st->p.seed, sid = st->p.stream_id;
    double vol = 1.0;
    if (st->p.has_domain)
        for (int i = 0; i < d; ++i) vol *= st->domain_host[d + i];
    void* args[] = {(void*)&map,   (void*)&st->nh_sum, (void*)&C,     (void*)&n,    (void*)&itu,
                    (void*)&st->domain, (void*)&vol,   (void*)&st->theta, (void*)&seed, (void*)&sid,
                    (void*)&d_cube, (void*)&d_ids,     (void*)&d_x,   (void*)&d_jf};
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipError_t e = hipModuleLaunchKernel(st->prog->dump, grid, 1, 1, 256, 1, 1, 0, st->stream, args, nullptr);
    if (e == hipSuccess) e = hipMemcpyAsync(cube, d_cube, sizeof(long long) * n, hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ids, d_ids, sizeof(int) * n * d, hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(x, d_x, sizeof(double) * n * d, hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(jf, d_jf, sizeof(double) * n, hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    (void)hipFree(d_cube); (void)hipFree(d_ids); (void)hipFree(d_x); (void)hipFree(d_jf);
    if (e != hipSuccess) return fail(CVG_ERR_HIP, std::string("dump: ") + hipGetErrorString(e));
    return CVG_OK;
}


int cvg_state_set_record(cvg_state* st, int enable) {
    if (!st) return fail(CVG_ERR_INVALID, "state is NULL");
    if (enable && !st->snap_x) {
        const long long slots = st->p.max_it - st->p.skip;
        if (slots < 1) return fail(CVG_ERR_INVALID, "nothing to record: skip >= max_it");
        const long long nx = (long long)st->p.n_dim * (st->p.n_intervals + 1);
        int rc;
        if ((rc = dalloc(&st->snap_x, slots * nx))) return rc;
        if ((rc = dalloc(&st->snap_nh, slots * st->C))) return rc;
    }
    if ((enable ? 1 : 0) != st->record) {
        if (st->graph_exec) { (void)hipGraphExecDestroy(st->graph_exec); st->graph_exec = nullptr; }
        if (st->graph) { (void)hipGraphDestroy(st->graph); st->graph = nullptr; }
    }
    st->record = enable ? 1 : 0;
    return CVG_OK;
}

int cvg_state_get_snapshot(cvg_state* st, int it, double* x_edges, int* nh) {
    if (!st || !x_edges || !nh) return fail(CVG_ERR_INVALID, "NULL argument");
    if (it < 0 || it >= st->iterations_done || !st->recorded[it])
        return fail(CVG_ERR_INVALID, "iteration was not recorded");
    const long long slot = it - st->p.skip;
    const long long nx = (long long)st->p.n_dim * (st->p.n_intervals + 1);
    int rc = copy_d2h(st, x_edges, st->snap_x + slot * nx, sizeof(double) * nx);
    if (rc) return rc;
    return copy_d2h(st, nh, st->snap_nh + slot * st->C, sizeof(int) * st->C);
}

int cvg_state_backward(cvg_state* st, const double* theta, double* grad, double* grad_var,
                       double* value, double* value_var) {
    vg_bw_layout L;
    int rc = vg_backward_layout(st, &L);
    if (rc) return rc;
    if (!st->bw_part) {
        // the state's own replay buffers, allocated by the first call and
        // reused by every later one
        if ((rc = dalloc(&st->bw_part, L.part))) return rc;
        if ((rc = dalloc((double**)&st->bw_theta, L.theta))) return rc;
        if ((rc = dalloc(&st->bw_nh_sum, L.nh_sum))) return rc;
        if ((rc = dalloc(&st->bw_s_total, 1))) return rc;
        if ((rc = dalloc(&st->bw_x, L.x))) return rc;
        if ((rc = dalloc(&st->bw_map64, L.map))) return rc;
        if ((rc = dalloc(&st->bw_map32, L.map))) return rc;
        if ((rc = dalloc(&st->bw_fold_list, 1))) return rc;
        if ((rc = dalloc(&st->bw_fold_count, 2))) return rc;
        if ((rc = dalloc(&st->bw_keys, 2 * L.cubes))) return rc;
        if ((rc = dalloc(&st->bw_idx, 4 * L.cubes + 2 + L.waves + 1))) return rc;
        if ((rc = dalloc((char**)&st->bw_sort_tmp, L.sort_tmp))) return rc;
        st->bw_evs.assign(2 * (size_t)(st->p.max_it - st->p.skip), nullptr);
        for (hipEvent_t& e : st->bw_evs) VG_HIP(hipEventCreate(&e));
    }
    vg_bw_buffers b;
    b.part = st->bw_part;
    b.theta = st->bw_theta;
    b.nh_sum = st->bw_nh_sum;
    b.s_total = st->bw_s_total;
    b.x = st->bw_x;
    b.map64 = st->bw_map64;
    b.map32 = st->bw_map32;
    b.fold_list = st->bw_fold_list;
    b.fold_count = st->bw_fold_count;
    b.keys = st->bw_keys;
    b.keys_sorted = st->bw_keys + L.cubes;
    b.idx = st->bw_idx;
    b.order = st->bw_idx + L.cubes;
    b.lanes = st->bw_idx + 2 * L.cubes;
    b.lane_start = st->bw_idx + 3 * L.cubes + 1;
    b.wave_first = st->bw_idx + 4 * L.cubes + 2;
    b.sort_tmp = st->bw_sort_tmp;
    return vg_backward_run(st, theta, b, L, st->bw_evs.data(), grad, grad_var, value, value_var);
}

int cvg_state_backward_timing(cvg_state* st, double* replay_ms) {
    if (!st || !replay_ms) return fail(CVG_ERR_INVALID, "NULL argument");
    *replay_ms = st->bw_replay_ms;
    return CVG_OK;
}

}  // extern "C"

namespace {
// Programs compiled by cvg_integrate_cuda, kept for later calls with the same
// integrand source, compile-time shape, options and device (most recently
// used first, at most kOneShotPrograms).  The one-shot call is the only user;
// a program in use stays alive through its shared_ptr when it is evicted.
constexpr size_t kOneShotPrograms = 16;
struct one_shot_cache {
    std::mutex mu;
    std::list<std::pair<std::string, std::shared_ptr<cvg_program>>> lru;
};
one_shot_cache& one_shot_programs() {
    static one_shot_cache c;
    return c;
}

int one_shot_program(const char* src, const cvg_params& p, const std::string& opts,
                     std::shared_ptr<cvg_program>* out) {
    int dev = 0;
    VG_HIP(hipGetDevice(&dev));
    const std::string key = std::string(src) + '\x01' + std::to_string(dev) + ',' + std::to_string(p.n_dim) + ',' +
                            std::to_string(p.n_intervals) + ',' + std::to_string(p.n_strat) + ',' +
                            std::to_string(p.precision) + ',' + opts;
    one_shot_cache& c = one_shot_programs();
    std::lock_guard<std::mutex> lock(c.mu);
    for (auto it = c.lru.begin(); it != c.lru.end(); ++it) {
        if (it->first == key) {
            c.lru.splice(c.lru.begin(), c.lru, it);
            *out = it->second;
            return CVG_OK;
        }
    }
    cvg_program* raw = nullptr;
    int rc = cvg_program_create(src, &p, opts.c_str(), &raw);
    if (!rc) rc = raw->load();  // loaded here, under the lock, on this device
    if (rc) {
        cvg_program_destroy(raw);
        return rc;
    }
    out->reset(raw, cvg_program_destroy);
    c.lru.emplace_front(key, *out);
    if (c.lru.size() > kOneShotPrograms) c.lru.pop_back();
    return CVG_OK;
}
}  // namespace

extern "C" {

int cvg_integrate_cuda(long long tot_eval, int max_it, int skip, long long max_batch_size,
                       int n_intervals, int n_strat, int n_dim, double alpha, double beta,
                       unsigned int seed, int multi_gpu, int multi_cpu, int gpu_num, int cpu_num,
                       double gpu_thr, const char* integrand_src, const char* options,
                       cvg_ref_result* out) {
    (void)max_batch_size;
    (void)cpu_num;
    (void)gpu_num;
    const double t0 = now_ms();
    if (!out) return fail(CVG_ERR_INVALID, "out is NULL");
    if (max_it < 1) return fail(CVG_ERR_INVALID, "max_it must be >= 1");
    cvg_params p{};
    p.n_dim = n_dim;
    p.n_intervals = n_intervals;
    p.n_strat = n_strat;
    p.max_it = max_it;
    p.skip = skip;
    p.n_eval_it = tot_eval / max_it;  // cuda_wrapper.cu:331
    p.alpha = alpha;
    p.beta = beta;
    p.seed = seed;
    p.precision = CVG_FP64;
    std::string rtc_opts;
    if (options) {
        std::istringstream ss(options);
        std::string tok;
        while (ss >> tok) {
            if (tok == "fp32") p.precision = CVG_FP32C;
            else if (tok == "fp64") p.precision = CVG_FP64;
            else rtc_opts += tok + " ";
        }
    }
    if (!integrand_src) return fail(CVG_ERR_INVALID, "integrand_src is NULL");
    std::shared_ptr<cvg_program> prog;
    int rc = one_shot_program(integrand_src, p, rtc_opts, &prog);
    if (rc) return rc;
    cvg_state* st = nullptr;
    rc = cvg_state_create(prog.get(), &p, &st);
    if (!rc) rc = cvg_state_run(st, max_it, 1);
    cvg_result r{};
    if (!rc) rc = cvg_state_result(st, &r);
    cvg_state_destroy(st);
    if (rc) return rc;
    out->res = r.result;
    out->err = r.error;
    out->n_strat = n_strat;
    out->n_cubes = r.n_cubes;
    out->n_intervals = n_intervals;
    out->tot_nevals = r.tot_nevals;
    out->tot_time = now_ms() - t0;
    out->avg_it_time = r.loop_ms / max_it;
    out->seed = seed;
    out->multi_gpu = multi_gpu;
    out->multi_cpu = multi_cpu;
    out->gpu_num = 1;
    out->omp_threads_num = 0;
    out->gpu_thr = gpu_thr;
    return CVG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: many independent integrals per launch (no reference counterpart).
// ---------------------------------------------------------------------------

// Item descriptor of the batched sampling kernel (layout of vg_item in vg_fill.hip).
struct vg_fitem {
    long long cube_off, nhs_off, win_off;
    double V, inv_ns;
    int d, C, map_off, theta_off;
    unsigned ns, ns_m, ns_sh1, ns_sh2;
    float inv_ns_f, pad;
};
static_assert(sizeof(vg_fitem) == 80, "vg_fitem layout");

// Magic numbers for floor(x / d), 32-bit x (Granlund & Montgomery 1994, fig. 4.1).
static void udiv_magic(unsigned d, unsigned* m, unsigned* sh1, unsigned* sh2) {
    int l = 0;
    while ((1ull << l) < d) ++l;
    *m = (unsigned)((((1ull << l) - d) << 32) / d + 1);
    *sh1 = l < 1 ? (unsigned)l : 1u;
    *sh2 = l > 1 ? (unsigned)(l - 1) : 0u;
}

template <typename T>
static int upload(hipStream_t st, T** dptr, const std::vector<T>& v) {
    int rc = dalloc(dptr, (long long)v.size());
    if (rc) return rc;
    if (!v.empty()) {
        VG_HIP(hipMemcpyAsync(*dptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
    }
    return CVG_OK;
}

struct cvg_batch {
    cvg_program* prog = nullptr;
    cvg_params p{};
    int n_items = 0, K = kDefaultRun, k_shift = 4, device = -1;
    long long tot_cubes = 0, tot_nhs = 0, tot_win = 0, tot_blk = 0, tot_xe = 0, tot_map = 0;
    int n_refine = 0, n_tasks = 0;
    std::vector<vg_bitem> hitems;
    std::vector<int> dims;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    vg_fitem* fitems = nullptr;
    vg_bitem* bitems = nullptr;
    int *task_item = nullptr, *task_w0 = nullptr, *blk_item = nullptr, *rd_item = nullptr, *rd_dim = nullptr;
    int* nh = nullptr;
    long long* nh_sum = nullptr;
    double2* JS = nullptr;  // per-cube (sum J f, sum (J f)^2)
    double* dh = nullptr;
    int* run_cube = nullptr;
    double2 *win_first = nullptr, *win_last = nullptr;
    double *part = nullptr, *dsum = nullptr;
    long long *bsum = nullptr, *boff = nullptr, *s_item = nullptr;
    int *it_p = nullptr, *err = nullptr;
    double* hist = nullptr;
    double* x_all = nullptr;
    double2* map64 = nullptr;
    float2* map32 = nullptr;
    double* weights = nullptr;
    unsigned* counts = nullptr;
    void* theta = nullptr;
    long long* fold_list = nullptr;
    int* fold_count = nullptr;  // [0] fold list length, [1] big-run list length
    int fold_cap = 0;
    long long* big_list = nullptr;  // run-table chunks of cubes holding many runs
    int big_cap = 0;
    int iterations_done = 0;
    double loop_ms = 0.0, setup_ms = 0.0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // optional per-iteration timing (eager launches): sampling launch and
    // whole iteration, as for a single integral
    int timing = 0;
    std::vector<hipEvent_t> tev;
    std::vector<double> fill_ms, iter_ms;

    ~cvg_batch() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (hipEvent_t e : tev) (void)hipEventDestroy(e);
        void* bufs[] = {fitems, bitems, task_item, task_w0, blk_item, rd_item, rd_dim, nh, nh_sum, JS,
                        dh, run_cube, win_first, win_last, part, dsum, bsum, boff, s_item, it_p, err,
                        hist, x_all, map64, map32, weights, counts, theta, fold_list, fold_count, big_list};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }

    // phase 0: initial map and allocation; 1: refinement, results, reallocation.
    // The refinement of phase 1 runs as extra blocks of the results launch.
    int steps(int phase) {
        if (phase == 0)
            VG_HIP(vg_launch_batch_refine(stream, bitems, n_refine, rd_item, rd_dim, p.n_intervals, p.alpha, 1,
                                          x_all, weights, counts, map64, map32));
        VG_HIP(vg_launch_batch_iteration_steps(
            stream, bitems, n_items, tot_blk, blk_item, n_refine, rd_item, rd_dim, p.n_intervals, p.alpha,
            p.beta, k_shift, p.max_it, nh, nh_sum, JS, win_first, win_last, dh, part, dsum, bsum, boff,
            s_item, run_cube, it_p, hist, x_all, weights, counts, map64, map32, err, phase, fold_list,
            fold_count, fold_cap, task_item, task_w0, n_tasks, (prog->block / 64) * prog->task_w, big_list,
            big_cap));
        return CVG_OK;
    }

    int fill() {
        int ncu = 0;
        VG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
        // fill_grid 1: one block per task, so the hardware hands tasks (whose
        // cost varies with the item's dimension and integrand) to CUs as they
        // free up; 0: a resident grid striding over the task list
        unsigned grid = (unsigned)(ncu * prog->fill_blocks_per_cu);
        if (prog->fill_grid || (long long)grid > n_tasks) grid = (unsigned)n_tasks;
        const void* map = p.precision == CVG_FP32C ? (const void*)map32 : (const void*)map64;
        unsigned seed = p.seed;
        int nt = n_tasks;
        void* args[] = {(void*)&fitems,  (void*)&task_item, (void*)&task_w0,  (void*)&nt,
                        (void*)&map,     (void*)&nh_sum,    (void*)&run_cube, (void*)&s_item,
                        (void*)&it_p,    (void*)&JS,        (void*)&win_first, (void*)&win_last,
                        (void*)&weights, (void*)&counts,    (void*)&theta,     (void*)&seed};
        VG_HIP(hipModuleLaunchKernel(prog->bfill, grid, 1, 1, (unsigned)prog->block, 1, 1, 0, stream, args,
                                     nullptr));
        return CVG_OK;
    }

    int iteration(hipEvent_t e_start = nullptr, hipEvent_t e_fill = nullptr, hipEvent_t e_end = nullptr) {
        if (e_start) VG_HIP(hipEventRecord(e_start, stream));
        int rc = fill();
        if (rc) return rc;
        if (e_fill) VG_HIP(hipEventRecord(e_fill, stream));
        if ((rc = steps(1))) return rc;
        if (e_end) VG_HIP(hipEventRecord(e_end, stream));
        return CVG_OK;
    }
};

extern "C" {

int cvg_batch_program_create(const char* integrand_src, int max_dim, const cvg_params* shape,
                             const char* options, cvg_program** out) {
    if (!out) return fail(CVG_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!integrand_src || !shape) return fail(CVG_ERR_INVALID, "NULL argument");
    if (max_dim < 1 || max_dim > 16) return fail(CVG_ERR_INVALID, "max_dim must be in 1..16");
    if (shape->n_intervals < 2 || shape->n_intervals > 1024)
        return fail(CVG_ERR_INVALID, "batched n_intervals must be in 2..1024");
    if (shape->precision != CVG_FP64 && shape->precision != CVG_FP32C)
        return fail(CVG_ERR_INVALID, "precision must be CVG_FP64 or CVG_FP32C");
    const int K = shape->run_len > 0 ? shape->run_len : kDefaultRun;
    if (K > 4096 || (K & (K - 1))) return fail(CVG_ERR_INVALID, "run_len must be a power of two up to 4096");
    if ((long long)max_dim * shape->n_intervals * 12 > kLdsBytes)
        return fail(CVG_ERR_INVALID, "max_dim * n_intervals bins do not fit in LDS");
    auto prog = std::make_unique<cvg_program>();
    prog->batch = 1;
    prog->max_dim = max_dim;
    prog->d = max_dim;
    prog->nbins = shape->n_intervals;
    prog->precision = shape->precision;
    prog->K = K;
    int block = kFillBlock, bcopies = 0;
    std::string rtc_extra;
    if (options) {
        std::istringstream ss(options);
        std::string tok;
        while (ss >> tok) {
            const size_t eq = tok.find('=');
            const std::string key = eq == std::string::npos ? tok : tok.substr(0, eq);
            const int val = eq == std::string::npos ? 1 : std::atoi(tok.c_str() + eq + 1);
            if (key == "block") {
                if (val < 64 || val > 1024 || val % 64) return fail(CVG_ERR_INVALID, "block must be 64..1024, multiple of 64");
                block = val;
            } else if (key == "bin_copies") {
                if (val != 1 && val != 2 && val != 4) return fail(CVG_ERR_INVALID, "bin_copies must be 1, 2 or 4");
                bcopies = val;
            } else if (key == "task_windows") {
                if (val < 1 || val > 64) return fail(CVG_ERR_INVALID, "task_windows must be in 1..64");
                prog->task_w = val;
            } else if (key == "fill_grid") {
                if (val < 0 || val > 1) return fail(CVG_ERR_INVALID, "fill_grid must be 0 or 1");
                prog->fill_grid = val;
            } else {
                rtc_extra += tok + " ";
            }
        }
    }
    prog->block = block;
    // Copies of the block's LDS bins (lane l adds into copy l mod copies): the
    // 64 lanes of a wave meet in one of 256 bins often, and every block on a
    // CU adds into its own bins, so copies pay where they cost no blocks per
    // CU.  Default: 2 copies when the blocks that the CU's 16 waves make
    // (16 / waves per block) still fit in LDS with them, else 1.  512-integral
    // batch, 256 bins, max_dim 12: block 512 with 2 copies (72 KB) 92 ms for
    // 10 iterations against 96.7 ms for block 256 with 1 (36 KB; 108 ms with
    // 2, two blocks per CU) and 121 ms for block 512 with 1.
    const long long bin_lds = (long long)max_dim * prog->nbins * 12;
    auto copies_lds = [&](int c) { return c * bin_lds + (c - 1) * 17LL * 12; };
    if (bcopies == 0) {
        const int waves = block / 64;
        const long long per_cu = waves >= 16 ? 1 : 16 / waves;
        bcopies = copies_lds(2) * per_cu <= kLdsBytes ? 2 : 1;
    }
    if (copies_lds(bcopies) > kLdsBytes) return fail(CVG_ERR_INVALID, "the bin copies do not fit in LDS");
    prog->bin_copies = bcopies;
    const double t0 = now_ms();
    hiprtcProgram rtc;
    const char* hdr_src[3] = {kVgPhiloxSrc, kVgMathSrc, integrand_src};
    const char* hdr_name[3] = {"vg_philox.h", "vg_math.h", "vg_user.h"};
    if (hiprtcCreateProgram(&rtc, kVgFillSrc, "vg_fill.hip", 3, hdr_src, hdr_name) != HIPRTC_SUCCESS)
        return fail(CVG_ERR_COMPILE, "hiprtcCreateProgram failed");
    std::vector<std::string> opts = {
        "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-DVG_BATCH=1", "-DVG_MAXD=" + std::to_string(max_dim),
        "-DVG_BCOPIES=" + std::to_string(bcopies),
        "-DVG_NDIM=" + std::to_string(max_dim), "-DVG_NBINS=" + std::to_string(prog->nbins),
        "-DVG_NSTRAT=1u", "-DVG_K=" + std::to_string(K), "-DVG_BLOCK=" + std::to_string(block),
        "-DVG_FP32=" + std::to_string(prog->precision == CVG_FP32C ? 1 : 0),
        "-DVG_TASKW=" + std::to_string(prog->task_w)};
    {
        std::istringstream ss(rtc_extra);
        std::string tok;
        while (ss >> tok) opts.push_back(tok);
    }
    std::vector<const char*> copts;
    for (auto& o : opts) copts.push_back(o.c_str());
    hiprtcResult cr = hiprtcCompileProgram(rtc, (int)copts.size(), copts.data());
    size_t log_size = 0;
    hiprtcGetProgramLogSize(rtc, &log_size);
    if (log_size > 1) {
        prog->log.resize(log_size);
        hiprtcGetProgramLog(rtc, &prog->log[0]);
        while (!prog->log.empty() && prog->log.back() == '\0') prog->log.pop_back();
    }
    if (cr != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&rtc);
        *out = prog.release();
        return fail(CVG_ERR_COMPILE, std::string("hipRTC compile failed: ") + hiprtcGetErrorString(cr) + "\n" +
                                         (*out)->log);
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(rtc, &code_size);
    prog->code.resize(code_size);
    hiprtcGetCode(rtc, prog->code.data());
    hiprtcDestroyProgram(&rtc);
    prog->compile_ms = now_ms() - t0;
    *out = prog.release();
    return CVG_OK;
}

int cvg_batch_create(cvg_program* prog, const cvg_params* common, const cvg_batch_item* items, int n_items,
                     const double* theta, int n_theta, cvg_batch** out) {
    if (!out) return fail(CVG_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!prog || !prog->batch || prog->code.empty()) return fail(CVG_ERR_INVALID, "not a batch program");
    if (!common || !items || n_items < 1) return fail(CVG_ERR_INVALID, "no items");
    if (common->n_intervals != prog->nbins || common->precision != prog->precision ||
        (common->run_len > 0 ? common->run_len : kDefaultRun) != prog->K)
        return fail(CVG_ERR_INVALID, "common params do not match the batch program");
    if (common->max_it < 1) return fail(CVG_ERR_INVALID, "max_it must be >= 1");
    if (n_theta < 0 || (n_theta > 0 && !theta)) return fail(CVG_ERR_INVALID, "bad theta");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(CVG_ERR_NODEVICE, "no HIP device visible");
    int rc;
    if ((rc = prog->load())) return rc;
    const double t0 = now_ms();
    auto b = std::make_unique<cvg_batch>();
    b->prog = prog;
    b->p = *common;
    b->p.domain = nullptr;
    b->p.has_domain = 0;
    b->n_items = n_items;
    b->K = prog->K;
    b->k_shift = 0;
    while ((1 << b->k_shift) < b->K) ++b->k_shift;
    VG_HIP(hipGetDevice(&b->device));
    const int n = common->n_intervals;
    const long long WS = 64LL * b->K;
    const int WPB = prog->block / 64;
    std::vector<vg_fitem> fi((size_t)n_items);
    std::vector<vg_bitem>& bi = b->hitems;
    bi.resize((size_t)n_items);
    std::vector<int> task_item, task_w0, blk_item, rd_item, rd_dim;
    for (int i = 0; i < n_items; ++i) {
        const cvg_batch_item& it = items[i];
        if (it.n_dim < 1 || it.n_dim > prog->max_dim) return fail(CVG_ERR_INVALID, "item dimension out of range");
        if (it.n_eval_it < 1) return fail(CVG_ERR_INVALID, "item budget must be >= 1");
        if (it.theta_offset < 0 || (n_theta > 0 ? it.theta_offset >= n_theta : it.theta_offset != 0))
            return fail(CVG_ERR_INVALID, "item theta_offset out of range");
        int ns = it.n_strat;
        if (ns <= 0) ns = (int)std::pow(it.n_eval_it / 4.0, 1.0 / it.n_dim);  // python_wrapper.py:15-18
        if (ns < 1) ns = 1;
        if (ns >= (1 << 16)) return fail(CVG_ERR_INVALID, "item n_strat must be < 65536");
        long long C;
        if (!checked_cubes(ns, it.n_dim, &C)) return fail(CVG_ERR_INVALID, "item n_strat^n_dim must be < 2^31");
        const double V = std::pow(1.0 / ns, it.n_dim);
        const double v0 = 1.0 / (double)C * (double)it.n_eval_it;
        int nh0 = v0 != v0 ? 0 : (v0 >= 2147483647.0 ? 2147483647 : (int)v0);
        if (nh0 < 2) nh0 = 2;
        long long s_cap = it.n_eval_it + 2 * C + WS + 1024;
        if ((long long)nh0 * C > s_cap) s_cap = (long long)nh0 * C;
        const long long wcap = (s_cap + WS - 1) / WS + 1;
        const long long nblk = (C + 1023) / 1024;
        vg_fitem& f = fi[(size_t)i];
        vg_bitem& g = bi[(size_t)i];
        f.cube_off = g.cube_off = b->tot_cubes;
        f.nhs_off = g.nhs_off = b->tot_nhs;
        f.win_off = g.win_off = b->tot_win;
        g.wcap = wcap;
        g.s_cap = s_cap;
        g.blk_off = b->tot_blk;
        g.nblk = nblk;
        g.xe_off = b->tot_xe;
        f.map_off = (int)b->tot_map;
        g.map_off = b->tot_map;
        g.N = it.n_eval_it;
        f.V = g.V = V;
        f.inv_ns = 1.0 / (double)ns;
        f.inv_ns_f = 1.0f / (float)ns;
        f.pad = 0.0f;
        f.d = g.d = it.n_dim;
        f.C = g.C = (int)C;
        g.nh0 = nh0;
        g.pad = 0;
        f.theta_off = it.theta_offset;
        f.ns = (unsigned)ns;
        udiv_magic((unsigned)ns, &f.ns_m, &f.ns_sh1, &f.ns_sh2);
        const long long task_windows = (long long)WPB * prog->task_w;
        for (long long t = 0; t * task_windows < wcap; ++t) {
            task_item.push_back(i);
            task_w0.push_back((int)(t * task_windows));
        }
        for (long long k = 0; k < nblk; ++k) blk_item.push_back(i);
        for (int dd = 0; dd < it.n_dim; ++dd) { rd_item.push_back(i); rd_dim.push_back(dd); }
        b->dims.push_back(it.n_dim);
        // item ranges padded so every item's n_h starts on a 16-byte and its prefix
        // sums on a 16-byte boundary (vector accesses in the scan kernel)
        b->tot_cubes += (C + 3) / 4 * 4;
        b->tot_nhs += (C + 2) / 2 * 2;
        b->tot_win += wcap;
        b->tot_blk += nblk;
        b->tot_xe += (long long)it.n_dim * (n + 1);
        b->tot_map += (long long)it.n_dim * n;
        if (b->tot_win * 64 > 2147483647LL * 16) return fail(CVG_ERR_CAPACITY, "batch too large");
        // the sampling kernel indexes maps and bins with 32-bit (dim, bin) offsets
        if (b->tot_map > 2147483647LL) return fail(CVG_ERR_CAPACITY, "batch too large: more than 2^31 map entries");
    }
    b->n_tasks = (int)task_item.size();
    b->n_refine = (int)rd_item.size();
    VG_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->own_stream = true;
    VG_HIP(hipEventCreate(&b->ev0));
    VG_HIP(hipEventCreate(&b->ev1));
    hipStream_t s = b->stream;
    if ((rc = upload(s, &b->fitems, fi))) return rc;
    if ((rc = upload(s, &b->bitems, bi))) return rc;
    if ((rc = upload(s, &b->task_item, task_item))) return rc;
    if ((rc = upload(s, &b->task_w0, task_w0))) return rc;
    if ((rc = upload(s, &b->blk_item, blk_item))) return rc;
    if ((rc = upload(s, &b->rd_item, rd_item))) return rc;
    if ((rc = upload(s, &b->rd_dim, rd_dim))) return rc;
    if ((rc = dalloc(&b->nh, b->tot_cubes))) return rc;
    if ((rc = dalloc(&b->nh_sum, b->tot_nhs))) return rc;
    if ((rc = dalloc(&b->JS, b->tot_cubes))) return rc;
    if ((rc = dalloc(&b->dh, b->tot_cubes))) return rc;
    if ((rc = dalloc(&b->run_cube, b->tot_win * 64))) return rc;
    if ((rc = dalloc(&b->win_first, b->tot_win))) return rc;
    if ((rc = dalloc(&b->win_last, b->tot_win))) return rc;
    if ((rc = dalloc(&b->part, (long long)VG_NRES * b->tot_blk))) return rc;
    if ((rc = dalloc(&b->bsum, b->tot_blk))) return rc;
    if ((rc = dalloc(&b->boff, b->tot_blk))) return rc;
    if ((rc = dalloc(&b->dsum, n_items))) return rc;
    if ((rc = dalloc(&b->s_item, n_items))) return rc;
    if ((rc = dalloc(&b->it_p, 1))) return rc;
    if ((rc = dalloc(&b->err, 1))) return rc;
    if ((rc = dalloc(&b->hist, (long long)VG_HIST * n_items * common->max_it))) return rc;
    if ((rc = dalloc(&b->x_all, b->tot_xe))) return rc;
    if ((rc = dalloc(&b->map64, b->tot_map))) return rc;
    if ((rc = dalloc(&b->map32, b->tot_map))) return rc;
    if ((rc = dalloc(&b->weights, b->tot_map))) return rc;
    if ((rc = dalloc(&b->counts, b->tot_map))) return rc;
    b->fold_cap = (int)(b->tot_win / 64 + 2 * n_items + 2);
    if ((rc = dalloc(&b->fold_list, b->fold_cap))) return rc;
    if ((rc = dalloc(&b->fold_count, 2))) return rc;
    b->big_cap = (int)vg_big_list_cap(b->tot_win * 64);
    if ((rc = dalloc(&b->big_list, 2LL * b->big_cap))) return rc;
    const int nth = n_theta > 0 ? n_theta : 1;
    if ((rc = dalloc((char**)&b->theta, 8LL * nth))) return rc;
    if (n_theta > 0) {
        if (prog->precision == CVG_FP32C) {
            std::vector<float> tf(theta, theta + n_theta);
            VG_HIP(hipMemcpyAsync(b->theta, tf.data(), 4 * (size_t)n_theta, hipMemcpyHostToDevice, s));
            VG_HIP(hipStreamSynchronize(s));
        } else {
            VG_HIP(hipMemcpyAsync(b->theta, theta, 8 * (size_t)n_theta, hipMemcpyHostToDevice, s));
        }
    } else {
        VG_HIP(hipMemsetAsync(b->theta, 0, 8, s));
    }
    VG_HIP(hipMemsetAsync(b->it_p, 0, sizeof(int), s));
    VG_HIP(hipMemsetAsync(b->err, 0, sizeof(int), s));
    if ((rc = b->steps(0))) return rc;
    VG_HIP(hipStreamSynchronize(s));
    b->setup_ms = now_ms() - t0;
    *out = b.release();
    return CVG_OK;
}

void cvg_batch_destroy(cvg_batch* b) {
    if (b && b->stream) (void)hipStreamSynchronize(b->stream);
    delete b;
}

int cvg_batch_run(cvg_batch* b, int n_iter, int use_graph) {
    if (!b) return fail(CVG_ERR_INVALID, "batch is NULL");
    if (n_iter < 0) return fail(CVG_ERR_INVALID, "n_iter must be >= 0");
    if (b->iterations_done + n_iter > b->p.max_it) return fail(CVG_ERR_CAPACITY, "more iterations than max_it");
    if (n_iter == 0) return CVG_OK;
    int rc;
    if (use_graph && !b->graph_exec) {
        VG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = b->iteration();
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ce != hipSuccess) return fail(CVG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        b->graph = g;
        VG_HIP(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    }
    const bool timed = b->timing && !use_graph;
    if (timed) {
        while ((int)b->tev.size() < 3 * n_iter) {
            hipEvent_t e;
            VG_HIP(hipEventCreate(&e));
            b->tev.push_back(e);
        }
    }
    VG_HIP(hipEventRecord(b->ev0, b->stream));
    for (int i = 0; i < n_iter; ++i) {
        if (use_graph) {
            VG_HIP(hipGraphLaunch(b->graph_exec, b->stream));
        } else if (timed) {
            if ((rc = b->iteration(b->tev[3 * i], b->tev[3 * i + 1], b->tev[3 * i + 2]))) return rc;
        } else if ((rc = b->iteration())) {
            return rc;
        }
    }
    VG_HIP(hipEventRecord(b->ev1, b->stream));
    VG_HIP(hipEventSynchronize(b->ev1));
    float ms = 0.0f;
    VG_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->loop_ms += ms;
    b->iterations_done += n_iter;
    if (timed) {
        for (int i = 0; i < n_iter; ++i) {
            float a = 0.0f, c = 0.0f;
            VG_HIP(hipEventElapsedTime(&a, b->tev[3 * i], b->tev[3 * i + 1]));
            VG_HIP(hipEventElapsedTime(&c, b->tev[3 * i], b->tev[3 * i + 2]));
            b->fill_ms.push_back(a);
            b->iter_ms.push_back(c);
        }
    }
    int h_err = 0;
    VG_HIP(hipMemcpy(&h_err, b->err, sizeof(int), hipMemcpyDeviceToHost));
    if (h_err & VG_ERR_CAPACITY) return fail(CVG_ERR_CAPACITY, "an item's allocation exceeded its buffers");
    if (h_err & VG_ERR_HIST) return fail(CVG_ERR_CAPACITY, "more iterations than max_it");
    return CVG_OK;
}

int cvg_batch_set_timing(cvg_batch* b, int enable) {
    if (!b) return fail(CVG_ERR_INVALID, "batch is NULL");
    b->timing = enable ? 1 : 0;
    return CVG_OK;
}

int cvg_batch_get_timing(cvg_batch* b, double* fill_ms, double* iter_ms, int cap) {
    if (!b) return fail(CVG_ERR_INVALID, "batch is NULL");
    const int n = (int)b->fill_ms.size();
    for (int i = 0; i < n && i < cap; ++i) {
        if (fill_ms) fill_ms[i] = b->fill_ms[i];
        if (iter_ms) iter_ms[i] = b->iter_ms[i];
    }
    return n;
}

int cvg_batch_history(cvg_batch* b, int item, double* I, double* sigma2, long long* nevals, int cap) {
    if (!b || item < 0 || item >= b->n_items) return fail(CVG_ERR_INVALID, "bad item");
    const int n = b->iterations_done;
    if (n == 0) return 0;
    VG_HIP(hipStreamSynchronize(b->stream));
    std::vector<double> h((size_t)VG_HIST * n);
    VG_HIP(hipMemcpy(h.data(), b->hist + (size_t)item * b->p.max_it * VG_HIST, sizeof(double) * h.size(),
                     hipMemcpyDeviceToHost));
    for (int i = 0; i < n && i < cap; ++i) {
        if (I) I[i] = h[VG_HIST * i];
        if (sigma2) sigma2[i] = h[VG_HIST * i + 1];
        if (nevals) nevals[i] = (long long)h[VG_HIST * i + 3];
    }
    return n;
}

int cvg_batch_results(cvg_batch* b, double* result, double* error, double* chi2_dof, long long* nevals,
                      double* loop_ms, double* error_unbiased, double* chi2_dof_unbiased) {
    if (!b) return fail(CVG_ERR_INVALID, "batch is NULL");
    VG_HIP(hipStreamSynchronize(b->stream));
    const int n = b->iterations_done, cap = b->p.max_it;
    std::vector<double> h((size_t)VG_HIST * b->n_items * cap);
    VG_HIP(hipMemcpy(h.data(), b->hist, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < b->n_items; ++i) {
        const vg_combined c = combine_record(h.data() + (size_t)i * cap * VG_HIST, n, b->p.skip);
        if (result) result[i] = c.result;
        if (error) error[i] = c.error;
        if (chi2_dof) chi2_dof[i] = c.chi2_dof;
        if (nevals) nevals[i] = c.tot;
        if (error_unbiased) error_unbiased[i] = c.error_unb;
        if (chi2_dof_unbiased) chi2_dof_unbiased[i] = c.chi2_dof_unb;
    }
    if (loop_ms) *loop_ms = b->loop_ms;
    return CVG_OK;
}

int cvg_batch_item_info(cvg_batch* b, int item, int* n_dim, int* n_strat, long long* n_cubes) {
    if (!b || item < 0 || item >= b->n_items) return fail(CVG_ERR_INVALID, "bad item");
    const vg_bitem& g = b->hitems[(size_t)item];
    if (n_dim) *n_dim = g.d;
    if (n_cubes) *n_cubes = g.C;
    if (n_strat) {
        int ns = 1;
        while (true) {
            long long c;
            if (!checked_cubes(ns + 1, g.d, &c) || c > g.C) break;
            ++ns;
        }
        *n_strat = ns;
    }
    return CVG_OK;
}

int cvg_batch_get_map(cvg_batch* b, int item, double* x_edges) {
    if (!b || !x_edges || item < 0 || item >= b->n_items) return fail(CVG_ERR_INVALID, "bad argument");
    const vg_bitem& g = b->hitems[(size_t)item];
    VG_HIP(hipStreamSynchronize(b->stream));
    VG_HIP(hipMemcpy(x_edges, b->x_all + g.xe_off, sizeof(double) * g.d * (b->p.n_intervals + 1),
                     hipMemcpyDeviceToHost));
    return CVG_OK;
}

int cvg_batch_get_nh(cvg_batch* b, int item, int* nh) {
    if (!b || !nh || item < 0 || item >= b->n_items) return fail(CVG_ERR_INVALID, "bad argument");
    const vg_bitem& g = b->hitems[(size_t)item];
    VG_HIP(hipStreamSynchronize(b->stream));
    VG_HIP(hipMemcpy(nh, b->nh + g.cube_off, sizeof(int) * g.C, hipMemcpyDeviceToHost));
    return CVG_OK;
}

int cvg_batch_get_cube_sums(cvg_batch* b, int item, double* jf, double* jf2) {
    if (!b || !jf || !jf2 || item < 0 || item >= b->n_items) return fail(CVG_ERR_INVALID, "bad argument");
    const vg_bitem& g = b->hitems[(size_t)item];
    VG_HIP(hipStreamSynchronize(b->stream));
    std::vector<double2> js((size_t)g.C);
    VG_HIP(hipMemcpy(js.data(), b->JS + g.cube_off, sizeof(double2) * g.C, hipMemcpyDeviceToHost));
    for (long long h = 0; h < g.C; ++h) { jf[h] = js[(size_t)h].x; jf2[h] = js[(size_t)h].y; }
    return CVG_OK;
}

}  // extern "C"
