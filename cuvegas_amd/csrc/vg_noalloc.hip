// Backward replay that allocates nothing: cvg_state_backward with its replay
// buffers in a static device array of the library instead of buffers the
// state allocates on its first call.  For callers that cannot allocate device
// memory (a training step captured in a HIP graph, a serving loop that must
// not touch the allocator).  No counterpart in the reference.
//
// The array (kNoallocBytes, 64 MiB) is part of the library's code object: each
// GPU the library runs on holds its own copy until the process ends.  Host code
// never uses the array's name as an address: the first call on a device looks
// up that device's copy with hipGetSymbolAddress (with the device current) and
// keeps the address; every later call on the device uses the kept address.
// Calls on one GPU share its copy, so calls that overlap on one GPU race; a
// call overwrites every part of the array it reads (nothing has to start at
// zero) and leaves the last call's contents behind.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/cuvegas.h"
#include "vg_internal.h"

namespace {

constexpr long long kNoallocBytes = 64LL << 20;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];  // device address of `scratch` on each device, once looked up

// Address of `scratch` on device dev, looked up once per device with dev current.
int scratch_address(int dev, void** out) {
    if (dev < 0 || dev >= kMaxDevices) return vg_fail(CVG_ERR_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        hipError_t r = hipSuccess;
        if (cur != dev) r = hipSetDevice(cur);
        if (e != hipSuccess)
            return vg_fail(CVG_ERR_HIP, (std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)).c_str());
        if (r != hipSuccess)
            return vg_fail(CVG_ERR_HIP, (std::string("hipSetDevice: ") + hipGetErrorString(r)).c_str());
        if (!p) return vg_fail(CVG_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return CVG_OK;
}

long long round256(long long bytes) { return (bytes + 255) & ~255LL; }

}  // namespace

extern "C" int cvg_state_backward_noalloc(cvg_state* st, const double* theta, double* grad, double* grad_var,
                                          double* value, double* value_var) {
    vg_bw_layout L;
    int rc = vg_backward_layout(st, &L);
    if (rc) return rc;
    // the buffers of vg_bw_buffers, each on a 256-byte boundary
    const long long sizes[17] = {round256(8 * L.part),  round256(8 * L.theta + 8), round256(8 * L.nh_sum),
                                 round256(8),           round256(8 * L.x),         round256(16 * L.map),
                                 round256(8 * L.map),   round256(4),               round256(8),
                                 round256(L.cubes),     round256(L.cubes),         round256(4 * L.cubes),
                                 round256(4 * L.cubes), round256(4 * L.cubes + 4), round256(4 * L.cubes + 4),
                                 round256(4 * L.waves + 4), round256(L.sort_tmp)};
    long long total = 0;
    for (long long b : sizes) total += b;
    if (total > kNoallocBytes)
        return vg_fail(CVG_ERR_CAPACITY,
                       ("backward_noalloc: the replay of this state needs " + std::to_string(total) +
                        " bytes of scratch; the static buffer holds " + std::to_string(kNoallocBytes) +
                        " (use cvg_state_backward)").c_str());
    void* base = nullptr;
    if ((rc = scratch_address(vg_state_device(st), &base))) return rc;
    char* p = static_cast<char*>(base);
    vg_bw_buffers b;
    b.part = reinterpret_cast<double*>(p);
    p += sizes[0];
    b.theta = p;
    p += sizes[1];
    b.nh_sum = reinterpret_cast<long long*>(p);
    p += sizes[2];
    b.s_total = reinterpret_cast<long long*>(p);
    p += sizes[3];
    b.x = reinterpret_cast<double*>(p);
    p += sizes[4];
    b.map64 = reinterpret_cast<double2*>(p);
    p += sizes[5];
    b.map32 = reinterpret_cast<float2*>(p);
    p += sizes[6];
    b.fold_list = reinterpret_cast<int*>(p);
    p += sizes[7];
    b.fold_count = reinterpret_cast<int*>(p);
    p += sizes[8];
    b.keys = reinterpret_cast<unsigned char*>(p);
    p += sizes[9];
    b.keys_sorted = reinterpret_cast<unsigned char*>(p);
    p += sizes[10];
    b.idx = reinterpret_cast<int*>(p);
    p += sizes[11];
    b.order = reinterpret_cast<int*>(p);
    p += sizes[12];
    b.lanes = reinterpret_cast<int*>(p);
    p += sizes[13];
    b.lane_start = reinterpret_cast<int*>(p);
    p += sizes[14];
    b.wave_first = reinterpret_cast<int*>(p);
    p += sizes[15];
    b.sort_tmp = p;
    return vg_backward_run(st, theta, b, L, nullptr, grad, grad_var, value, value_var);
}
