// Interface between the library's own source files (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

struct cvg_state;

// Device buffers of one backward replay.  Every call overwrites what it reads
// (the unit partials and totals, the frozen tables, the prefix sums, the
// counters, which the scan kernels clear themselves), so none of them has to
// start at zero.
struct vg_bw_buffers {
    double* part = nullptr;        // [units + max_it - skip][2 + 2 n_theta]: unit partials, then the
                                   // totals of each recorded iteration
    void* theta = nullptr;         // replay parameters (n_theta values of the working precision)
    long long* nh_sum = nullptr;   // [C + 1] prefix sums of a recorded allocation
    long long* s_total = nullptr;  // [1] samples of that allocation
    double* x = nullptr;           // [d (n + 1)] recorded map edges
    double2* map64 = nullptr;      // [d n] sampling tables of that map (FP64 form)
    float2* map32 = nullptr;       // [d n] (FP32 form)
    int* fold_list = nullptr;      // [1]
    int* fold_count = nullptr;     // [2]
    unsigned char* keys = nullptr;         // [C] size classes of the recorded allocation
    unsigned char* keys_sorted = nullptr;  // [C]
    int* idx = nullptr;                    // [C] cube indices
    int* order = nullptr;                  // [C] cubes by size, largest first
    int* lanes = nullptr;                  // [C + 1] lanes of each sorted cube
    int* lane_start = nullptr;             // [C + 1] their exclusive prefix sum
    int* wave_first = nullptr;             // [waves + 1] first sorted cube of each wave
    void* sort_tmp = nullptr;              // sort_tmp bytes of radix-sort and scan workspace
};

// Element counts of those buffers for a state, and the replay's launch shape.
struct vg_bw_layout {
    long long part = 0, theta = 0, nh_sum = 0, x = 0, map = 0, cubes = 0, waves = 0, sort_tmp = 0;
    int units = 0, slots_per_unit = 0;
};

// Checks that the state can be replayed (program compiled with the gradient
// kernel, recorded iterations) and gives the layout of its buffers.
int vg_backward_layout(const cvg_state* st, vg_bw_layout* out);

// Device the state's buffers and stream belong to.
int vg_state_device(const cvg_state* st);

// The replay of the recorded iterations through the buffers b (laid out by
// vg_backward_layout): value, gradient and their variances.  Every recorded
// iteration is enqueued on the state's stream, then one copy brings all their
// totals back.  evs, when not null, holds 2 (max_it - skip) events that time
// the replay kernels of each iteration (summed into the state's
// last-backward time).
int vg_backward_run(cvg_state* st, const double* theta, const vg_bw_buffers& b, const vg_bw_layout& L,
                    const hipEvent_t* evs, double* grad, double* grad_var, double* value, double* value_var);

// Sets the calling thread's last-error message (cvg_last_error) and returns code.
int vg_fail(int code, const char* msg);
