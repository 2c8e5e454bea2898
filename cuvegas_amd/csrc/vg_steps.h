// Launch helpers for the ahead-of-time step kernels (vg_steps.hip).
#pragma once
#include <hip/hip_runtime.h>

#define VG_ERR_HIST 1      // more iterations than the history buffer holds
#define VG_ERR_CAPACITY 2  // allocation exceeded the sample buffers

hipError_t vg_launch_results(hipStream_t st, int C, double V, double beta, const int* nh,
                             const long long* nh_sum, double* JF, double* JF2,
                             const double2* win_first, const double2* win_last, int ws_shift,
                             double* dh, double* part);
hipError_t vg_launch_finalize(hipStream_t st, long long nparts, double* part, const long long* s_total,
                              const int* it_p, int hist_cap, double* hist, double* dsum, int* err);
hipError_t vg_launch_alloc(hipStream_t st, int C, int mode, int nh_const, const double* dh,
                           const double* dsum, long long N, int* nh, long long* bsum);
hipError_t vg_launch_scan(hipStream_t st, int C, const int* nh, long long* bsum, long long* boff,
                          long long* s_total, long long s_cap, int* it_p, int advance, int k_shift,
                          long long run_cap, long long* nh_sum, int* run_cube, int* err);
hipError_t vg_launch_init_map(hipStream_t st, int d, int n, double* x_edges, double2* map64,
                              float2* map32, double* weights, unsigned* counts);
hipError_t vg_launch_refine(hipStream_t st, int d, int n, double alpha, double* x_edges,
                            double* weights, unsigned* counts, double* scratch, double2* map64,
                            float2* map32);
hipError_t vg_launch_tables(hipStream_t st, int d, int n, const double* x_edges, double2* map64,
                            float2* map32);
long long vg_refine_scratch_doubles(int d, int n);
hipError_t vg_launch_layout(hipStream_t st, int C, const int* nh, long long* bsum, long long* boff,
                            long long* s_total, long long s_cap, int* it_p, int k_shift, long long run_cap,
                            long long* nh_sum, int* run_cube, int* err);
hipError_t vg_launch_snapshot(hipStream_t st, const int* it_p, int skip, int cap, int C, int nx,
                              const double* x_edges, const int* nh, double* snap_x, int* snap_nh);
