// A standalone configuration in the reference's header format (settings as
// initialised globals, the integrand as a __host__ __device__ function of the
// point, src/config.cuh): the reference's default run, the 4-D ridge of 1000
// Gaussians, I = 0.851317758241297887.
#ifndef RIDGE4_CONFIG
#define RIDGE4_CONFIG
#include "commons/functions.cuh"
#define _USE_MATH_DEFINES
#include <math.h>

long int tot_eval = 10000000;   // samples over all iterations
int max_it = 20;
int skip = 5;                   // iterations left out of the average
long int max_batch_size = 1048576;

const int n_intervals = 1024;   // map bins per dimension
double alpha = 0.5;
double beta = 0.75;

const int blockSizeFill = 64;   // launch sizes of the reference's kernels: not used here

const int n_dim = 4;
__host__ __device__
double integrand(double* x) {
    return ridge(x, n_dim);
}

#endif
