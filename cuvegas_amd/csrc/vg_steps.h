// Launch helpers for the ahead-of-time step kernels (vg_steps.hip).
#pragma once
#include <hip/hip_runtime.h>

#define VG_ERR_HIST 1      // more iterations than the history buffer holds
#define VG_ERR_CAPACITY 2  // allocation exceeded the sample buffers

// Per-cube result sums reduced every iteration (block partials hold VG_NRES
// doubles per block): sum I_h, sum sigma_h^2 / n_h (the reference's variance),
// D = sum d_h, and sum sigma_h^2 / (n_h - 1) (the unbiased per-cube variance
// of the same estimate, reported beside the reference's).
#define VG_NRES 4
// Per-iteration record (doubles per iteration): I_k, sigma_k^2, D, samples,
// unbiased sigma_k^2.
#define VG_HIST 5

// Per-cube results from complete per-cube sums: vg_launch_bigfold must have
// folded the window partials of every cube that crosses a sampling window.
hipError_t vg_launch_results(hipStream_t st, int C, double V, double beta, const int* nh,
                             const double2* JS, double* dh, double* part);
// Results plus the map refinement of the d dimensions in d extra blocks of the
// same launch (needs vg_results_refine_lds(n) bytes of dynamic LDS per block).
hipError_t vg_launch_results_refine(hipStream_t st, int C, double V, double beta, const int* nh,
                                    const double2* JS, double* dh, double* part, int d, int n, double alpha,
                                    double* x_edges, double* weights, unsigned* counts, double* scratch,
                                    double2* map64, float2* map32);
long long vg_results_refine_lds(int n);
// part holds VG_NRES * (nparts + ceil(nparts / 1024)) doubles; sync is one zeroed
// counter that the launch leaves zeroed.
hipError_t vg_launch_finalize(hipStream_t st, long long nparts, double* part, unsigned* sync,
                              const long long* s_total, const int* it_p, int hist_cap, double* hist,
                              double* dsum, int* err);
hipError_t vg_launch_alloc(hipStream_t st, int C, int mode, int nh_const, const double* dh,
                           const double* dsum, long long N, int* nh, long long* bsum);
// bsum and boff hold nb + ceil(nb / 1024) int64 (nb = ceil(C / 1024)); sync is
// one zeroed counter that the launch leaves zeroed.  counts: 2 ints (lengths
// of the fold list and of big_list, 2 * big_cap int64 with big_cap =
// vg_big_list_cap(run_cap)).  run_cube may be NULL (no run table).
hipError_t vg_launch_scan(hipStream_t st, int C, const int* nh, long long* bsum, long long* boff,
                          unsigned* sync, long long* s_total, long long s_cap, int* it_p, int advance,
                          int k_shift, long long run_cap, long long* nh_sum, int* run_cube, int* err,
                          int* fold_list, int* counts, int fold_cap, long long* big_list, int big_cap);
long long vg_big_list_cap(long long run_cap);
hipError_t vg_launch_bigfold(hipStream_t st, const int* list, const int* count, int cap,
                             const long long* nh_sum, const int* run_cube, const long long* s_total,
                             const double2* win_first, const double2* win_last, int ws_shift,
                             double2* JS);
hipError_t vg_launch_init_map(hipStream_t st, int d, int n, double* x_edges, double2* map64,
                              float2* map32, double* weights, unsigned* counts);
hipError_t vg_launch_refine(hipStream_t st, int d, int n, double alpha, double* x_edges,
                            double* weights, unsigned* counts, double* scratch, double2* map64,
                            float2* map32);
hipError_t vg_launch_tables(hipStream_t st, int d, int n, const double* x_edges, double2* map64,
                            float2* map32);
long long vg_refine_scratch_doubles(int d, int n);
hipError_t vg_launch_layout(hipStream_t st, int C, const int* nh, long long* bsum, long long* boff,
                            unsigned* sync, long long* s_total, long long s_cap, int* it_p, int k_shift,
                            long long run_cap, long long* nh_sum, int* run_cube, int* err, int* fold_list,
                            int* counts, int fold_cap, long long* big_list, int big_cap);
hipError_t vg_launch_snapshot(hipStream_t st, const int* it_p, int skip, int cap, int C, int nx,
                              const double* x_edges, const int* nh, double* snap_x, int* snap_nh);

// Host-side descriptor of one item of a batch (the step kernels' view).
struct vg_bitem {
    long long cube_off;  // first cube of the item
    long long nhs_off;   // first nh_sum entry (C + 1 per item)
    long long win_off;   // first window
    long long wcap;      // window capacity (runs = 64 * wcap)
    long long s_cap;     // sample capacity
    long long blk_off;   // first 1024-cube block
    long long nblk;      // number of 1024-cube blocks
    long long xe_off;    // first map edge (d * (n + 1) per item)
    long long map_off;   // first (dim, bin) entry of maps and bins
    long long N;         // samples per iteration
    double V;            // cube volume
    int C, d, nh0, pad;
};

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
                                           int task_windows, long long* big_list, int big_cap);
hipError_t vg_launch_batch_refine(hipStream_t st, const vg_bitem* items, int n_refine, const int* rd_item,
                                  const int* rd_dim, int n, double alpha, int init, double* x_all,
                                  double* weights, unsigned* counts, double2* map64, float2* map32);

// Work layout of the backward replay (vg_grad) for the recorded allocation nh:
// order (C ints) the cubes sorted by sample count, largest first; lanes (C + 1
// ints) and lane_start (C + 1 ints, exclusive prefix) the lanes of each sorted
// cube; wave_first (vg_grad_max_waves(C, S) + 1 ints) the first sorted cube of
// each wave, then C.  keys, keys_sorted: C bytes; idx: C ints; tmp:
// vg_grad_order_tmp_bytes(C) bytes (0: the size query failed).  S bounds the
// allocation's samples.
size_t vg_grad_order_tmp_bytes(int C);
long long vg_grad_max_waves(long long C, long long S);
hipError_t vg_launch_grad_order(hipStream_t st, int C, const int* nh, unsigned char* keys,
                                unsigned char* keys_sorted, int* idx, int* order, int* lanes, int* lane_start,
                                int* wave_first, void* tmp, size_t tmp_bytes);
