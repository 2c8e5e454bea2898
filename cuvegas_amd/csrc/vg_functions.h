// The reference's integrand library (src/commons/functions.cuh:4-14, formulas
// of src/commons/functions.cu:5-111) with its signatures,
//   __host__ __device__ double name(double* x, int n_dim),
// so that a standalone configuration written for the reference
// (src/config.cuh: integrand(x) returning e.g. ridge(x, n_dim)) compiles
// unchanged under hipRTC: the standalone driver (vg_main.cpp) puts this file
// where the configuration includes "commons/functions.cuh".  Each function is
// the integrand's mathematical definition; the two that the reference
// evaluates partly in single precision (fabsf in roosarnold, sqrtf(12) in
// hellekalek) keep that, so results match the reference's program.  The same
// definitions, precision-generic, are the Python library
// (cuvegas_amd/integrands.py), where each has its closed form.
#pragma once

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

// sum_i x_i; integral d / 2.
__host__ __device__ inline double linear(double* x, int n_dim) {
    double s = 0.0;
    for (int i = 0; i < n_dim; ++i) s += x[i];
    return s;
}

// (1/d) sum_i |4 x_i - 2| (each term through the single-precision fabsf); integral 1.
__host__ __device__ inline double roosarnold(double* x, int n_dim) {
    double s = 0.0;
    for (int i = 0; i < n_dim; ++i) s += fabsf(4.0 * x[i] - 2.0);
    return s * (1.0 / (double)n_dim);
}

// (1 + 1/d)^d prod_i x_i^(1/d); integral 1.
__host__ __device__ inline double morokoffcalfisch(double* x, int n_dim) {
    const double e = 1.0 / (double)n_dim;
    double p = 1.0;
    for (int i = 0; i < n_dim; ++i) p *= pow(x[i], e);
    return p * pow(1.0 + e, n_dim);
}

// prod_i (x_i - 1/2) / sqrt(12) (the root in single precision); integral 0.
__host__ __device__ inline double hellekalek(double* x, int n_dim) {
    double p = 1.0;
    for (int i = 0; i < n_dim; ++i) p *= (x[i] - 0.5) / sqrtf(12.0f);
    return p;
}

// 1000 Gaussians exp(-100 |x - c_k|^2) centred at c_k = (k/999, ..., k/999),
// scaled by (100/pi)^2 / 1000; integral 0.851317758241297887 in 4-D.
__host__ __device__ inline double ridge(double* x, int n_dim) {
    const int n_peaks = 1000;
    const double scale = pow(100.0 / M_PI, 2.0) / n_peaks;
    double s = 0.0;
    for (int k = 0; k < n_peaks; ++k) {
        const double c = k / (n_peaks - 1.0);
        double r2 = 0.0;
        for (int i = 0; i < n_dim; ++i) r2 += (x[i] - c) * (x[i] - c);
        s += exp(-100.0 * r2);
    }
    return s * scale;
}

// exp(-|x|^2); integral (sqrt(pi)/2 erf 1)^d.
__host__ __device__ inline double expon(double* x, int n_dim) {
    double r2 = 0.0;
    for (int i = 0; i < n_dim; ++i) r2 += x[i] * x[i];
    return exp(-r2);
}

// prod_i cos x_i; integral (sin 1)^d.
__host__ __device__ inline double cosin(double* x, int n_dim) {
    double p = 1.0;
    for (int i = 0; i < n_dim; ++i) p *= cos(x[i]);
    return p;
}

// Dirichlet density with all parameters 5: prod_i x_i^4 / B, B = Gamma(5)^d / Gamma(5 d).
__host__ __device__ inline double dirichlet(double* x, int n_dim) {
    const double a = 5.0;
    double p = 1.0;
    for (int i = 0; i < n_dim; ++i) p *= pow(x[i], a - 1.0);
    return p / (pow(tgamma(a), n_dim) / tgamma(a * n_dim));
}

// Normalised Gaussian, mean 1/2 and variance 1e-4 in every coordinate; integral ~1.
__host__ __device__ inline double gaussian(double* x, int n_dim) {
    const double var = 0.0001;
    double r2 = 0.0;
    for (int i = 0; i < n_dim; ++i) r2 += (x[i] - 0.5) * (x[i] - 0.5);
    return exp(-0.5 * r2 / var) / pow(2.0 * M_PI * var, n_dim / 2.0);
}

// sin x_0 + exp x_1; integral (1 - cos 1) + (e - 1).
__host__ __device__ inline double sinexp(double* x, int n_dim) {
    (void)n_dim;
    return sin(x[0]) + exp(x[1]);
}

// 1; integral 1.
__host__ __device__ inline double one(double* x, int n_dim) {
    (void)x;
    (void)n_dim;
    return 1.0;
}
