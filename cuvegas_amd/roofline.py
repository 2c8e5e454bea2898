"""Algorithmic work per integrand evaluation and the VALU roofline of gfx950.

The sampling kernel has no dense contraction (no MFMA) and moves O(cubes)
bytes per iteration, not O(samples): it is bound by vector-ALU issue.  Work is
counted in VALU lane-slots per evaluation with fixed rules (one slot per VALU
operation, FP64 and FP32 alike since CDNA vector FP64 is full rate; a native
transcendental 4 slots plus 1 for range scaling; an LDS atomic 1 slot):

  S = 60 * ceil(d/4)   Philox4x32-10: 10 rounds x (2 mul_lo + 2 mul_hi + 2 xor3)
                       per 4 random words
    + 16 * d           uniform conversion 2, stratum/map/Jacobian 12, bin updates 2
    + S_f              the integrand
    + 10               f*J, (f*J)^2 and two compensated adds

with S_f = 2d+7 (Gaussian peak), 4d+14 (double Gaussian), d+6 (oscillatory),
d+9 (corner peak), K(2d+7) (K-component Gaussian mixture).

The backward replay of a parametrised integrand (d I / d theta, P
parameters) redraws each recorded sample and forms J f and the P terms
J df/dtheta_p: per replayed sample

  S_grad = S - 2 d     the forward's sample work without the 2 d bin updates
         + S_g         the integrand's gradient terms
         + 3 P         J df_p, and its sum and sum of squares

with S_g = K(5 + 2d) for the K-component Gaussian mixture (per component:
the weight term 1, the width term 4, 2 per centre coordinate).

Peak = CUs x 64 lanes x clock (lane-slots per second).

The same kernel against the LDS: every sample adds its (J f)^2 (ds_add_f64)
and a count (ds_add_u32) into one bin per dimension, at addresses as random as
the samples, so d such pairs per evaluation.  Their ceiling is measured on the
device (cvg_device_rates: CU cycles per wave64 instruction of the pair at
random slots of a 1024-slot array), and lds_roofline gives the fraction of the
launch the CUs' LDS would spend on the fill's pairs at that rate.  gfx950
issues some VALU operations (FP32 add/mul/fma, 32-bit add/shift/logic, moves)
in 2 cycles per wave64 instead of 4, so the slot peak above is a floor of the
VALU's capacity for FP32 code, not its ceiling; the LDS figure is the one that
bounds the FP32 sampling kernel.
"""
from __future__ import annotations

import math

INTEGRAND_SLOTS = {
    "gaussian_peak": lambda d, k=1: 2 * d + 7,
    "double_gaussian": lambda d, k=1: 4 * d + 14,
    "oscillatory": lambda d, k=1: d + 6,
    "corner_peak": lambda d, k=1: d + 9,
    "gaussian_mixture": lambda d, k=4: k * (2 * d + 7),
}


def slots_per_eval(d: int, family: str, components: int = 4) -> int:
    """VALU lane-slots per integrand evaluation (the rules above)."""
    s_f = INTEGRAND_SLOTS[family](d, components)
    return 60 * math.ceil(d / 4) + 16 * d + s_f + 10


GRADIENT_SLOTS = {
    "gaussian_mixture": lambda d, k=4: k * (5 + 2 * d),
}


def slots_per_replayed_sample(d: int, family: str, n_theta: int, components: int = 4) -> int:
    """VALU lane-slots per sample of the backward replay (the rules above)."""
    return (slots_per_eval(d, family, components) - 2 * d + GRADIENT_SLOTS[family](d, components)
            + 3 * n_theta)


def valu_peak(n_cu: int, clock_hz: float) -> float:
    """Lane-slots per second: every SIMD issues one 64-lane VALU op per cycle per CU."""
    return n_cu * 64 * clock_hz


RATE_NAMES = ("lds_f64_random", "lds_u32_random", "lds_pair_random", "lds_f64_seq", "lds_u32_seq",
              "valu_f32_fma", "valu_f64_fma", "valu_u32_add", "valu_mad_u64")


def device_rates() -> dict:
    """Measured cycles per wave64 instruction on the current device
    (cvg_device_rates), by name."""
    import ctypes as C

    from ._lib import check, lib
    buf = (C.c_double * len(RATE_NAMES))()
    check(lib.cvg_device_rates(buf, len(RATE_NAMES)))
    return {name: float(v) for name, v in zip(RATE_NAMES, buf)}


def lds_roofline(evals: float, n_dim: int, seconds: float, n_cu: int, clock_hz: float,
                 pair_cycles: float) -> dict:
    """LDS roofline of one sampling launch: evals * n_dim bin updates (a
    ds_add_f64 and a ds_add_u32 each) over the launch's duration, against the
    rate the CUs' LDS sustains for them at random addresses (pair_cycles: CU
    cycles per instruction of the pair, cvg_device_rates)."""
    pairs = evals * n_dim
    achieved = pairs / seconds
    peak = n_cu * clock_hz * 64.0 / (2.0 * pair_cycles)
    return {"bound": "lds", "achieved": achieved / 1e9, "peak": peak / 1e9, "unit": "G bin updates/s",
            "frac": achieved / peak, "updates_per_eval": n_dim, "cycles_per_instruction": pair_cycles}


MI355X_CU = 256
MI355X_CLOCK_HZ = 2.4e9
MI355X_HBM_BPS = 8.0e12
