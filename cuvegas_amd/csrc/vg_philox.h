// Philox4x32-10 counter-based generator and the sample -> (bin, offset) rule.
//
// Philox4x32-10 follows Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
// as easy as 1, 2, 3" (SC'11): ten rounds of two 32x32->64 multiplies with the
// round multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments
// 0x9E3779B9 / 0xBB67AE85.  It replaces the reference's per-thread cuRAND
// XORWOW state (vegas_cuda/src/vegas_cuda_extension/cuda_wrapper.cu:18-24,
// cuda_kernel.cu:21,39): a sample's random words are a pure function of its
// counter, so no generator state lives in memory and any sample can be
// regenerated (the autograd replay relies on that).
//
// Counter layout (fixed; every parity check depends on it):
//   ctr = { s mod 2^32, s div 2^32, iteration, group }   key = { seed, stream }
// where s is the sample's global index within the iteration (samples of cube h
// occupy [nh_sum[h], nh_sum[h] + nh[h])), group g yields the words for
// dimensions 4g .. 4g+3, and stream separates independent integrals.
//
// This header is compiled by hipRTC together with the integrand, and by hipcc
// for the ahead-of-time kernels.  It uses no library headers.
#pragma once

#ifndef VG_HD
#define VG_HD __host__ __device__ inline
#endif

struct vg_u4 { unsigned int x, y, z, w; };

// a ^ b ^ c.  On gfx950 one v_bitop3_b32 (truth table 0x96) instead of two
// v_xor_b32: the XORs are a third of the generator's instructions.
VG_HD unsigned int vg_xor3(unsigned int a, unsigned int b, unsigned int c) {
#if defined(__HIP_DEVICE_COMPILE__) && __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

// One Philox4x32 round with round key (k0, k1).
VG_HD vg_u4 vg_philox_round(vg_u4 c, unsigned int k0, unsigned int k1) {
    const unsigned long long p0 = (unsigned long long)0xD2511F53u * c.x;
    const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c.z;
    const unsigned int hi0 = (unsigned int)(p0 >> 32), lo0 = (unsigned int)p0;
    const unsigned int hi1 = (unsigned int)(p1 >> 32), lo1 = (unsigned int)p1;
    vg_u4 n;
    n.x = vg_xor3(hi1, c.y, k0);
    n.y = lo1;
    n.z = vg_xor3(hi0, c.w, k1);
    n.w = lo0;
    return n;
}

// Key of round r (0-based): the Weyl sequence k + r * (0x9E3779B9, 0xBB67AE85).
VG_HD unsigned int vg_philox_key0(unsigned int k0, int r) { return k0 + (unsigned int)r * 0x9E3779B9u; }
VG_HD unsigned int vg_philox_key1(unsigned int k1, int r) { return k1 + (unsigned int)r * 0xBB67AE85u; }

VG_HD vg_u4 vg_philox4x32_10(vg_u4 c, unsigned int k0, unsigned int k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round)
        c = vg_philox_round(c, vg_philox_key0(k0, round), vg_philox_key1(k1, round));
    return c;
}

// Random words for sample s, iteration it, dimension group g.
VG_HD vg_u4 vg_sample_words(unsigned long long s, unsigned int it, unsigned int g,
                            unsigned int seed, unsigned int stream) {
    vg_u4 c;
    c.x = (unsigned int)s;
    c.y = (unsigned int)(s >> 32);
    c.z = it;
    c.w = g;
    return vg_philox4x32_10(c, seed, stream);
}

VG_HD unsigned int vg_word(const vg_u4& v, int j) {
    return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
}

// Stratified point -> map bin, exactly in integers.
//
// The reference draws u in (0,1], forms y = (k + u) / n_strat for stratum digit
// k and takes bin id = (int)(y * n), offset = y * n - id
// (cuda_kernel.cu:35-46).  Here u = (2r + 1) / 2^33 for the 32-bit word r, so
// y * n = (k*n + (2r+1)*n / 2^33) / n_strat and
//   L   = (2r + 1) * n            (< 2^33 * n)
//   Q   = k*n + (L >> 33)
//   id  = Q / n_strat,  R = Q - id * n_strat
//   offset = (R + (L mod 2^33) * 2^-33) / n_strat   in [0, 1)
// which is exact integer arithmetic for the bin and identical on every device.
// Requires n_strat * n < 2^32 (checked on the host).
struct vg_binpos { unsigned int id; unsigned int rem; unsigned long long low; };

VG_HD vg_binpos vg_bin_of(unsigned int r, unsigned int kn, unsigned int nbins,
                          unsigned int nstrat) {
    const unsigned long long L = ((unsigned long long)r * 2ull + 1ull) * (unsigned long long)nbins;
    const unsigned int Q = kn + (unsigned int)(L >> 33);
    vg_binpos b;
    b.id = Q / nstrat;
    b.rem = Q - b.id * nstrat;
    b.low = L & ((1ull << 33) - 1ull);
    return b;
}
