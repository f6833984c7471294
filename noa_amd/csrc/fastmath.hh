// Table-driven FP64 log and exp (Tang's method) for the DCS kernels.
//
// ocml's FP64 log is ~98 VALU instructions and exp ~42 on gfx950, and the DCS
// path makes ~52 log and ~24 exp calls per (K, q) pair, so libm dominates the
// VALU issue budget. These versions take ~22 (log) and ~20 (exp) VALU
// instructions plus LDS table reads, within 0.75 ulp (log) and 0.6 ulp (exp).
// Special arguments are handled in place, without an inlined libm fallback at
// every call site: log of 0, negatives, inf and NaN by their IEEE values, of
// subnormals by scaling by 2^54; exp of NaN propagates, and the range beyond
// +-708 runs the same path, with ldexp producing the overflow, subnormal and
// zero results.
//
// log x = k ln2 + log(1/invc_i) + log1p(r),  r = z invc_i - 1 in [-2^-9, 2^-8],
//   with x = 2^k z, z in [0.5, 1) (x's mantissa under the exponent field
//   0x3fe: one and-or) and i from the top 8 mantissa bits (invc = 2 on the
//   first subinterval and 1 on the last, the two that hold x just above and
//   just below 1, so r = 2z - 1 or z - 1 exactly and log x keeps full
//   relative accuracy as x -> 1). The exponent enters as the biased field
//   E = k + 1022, with 1022 ln2 folded into the table: E LN2_HI + logc_hi is
//   exact (multiples of 2^-43 below 2^10), the small parts are summed before
//   they are added to it (exactly 0 at x -> 1: LN2_LO has 42 bits), and
//   log1p(r) - r = r^2 P(r) with P the degree-4 Chebyshev economization of
//   the series over that r range (truncation error < 0.07 * 2^-53 relative;
//   coefficients and tables from gen_fastmath_tables.py). Five VALU
//   instructions take x apart (two shifts, an and, a bitfield insert, a
//   conversion); a split around 0.6875, which centres the subintervals on 1
//   without special ones, takes seven.
// exp x = 2^m 2^(j/128) exp(r), k = rint(128 x / ln2) = 128 m + j,
//   r = x - k ln2/128 (two FMAs), exp(r) - 1 a degree-5 Taylor polynomial.
//
// The header is compiled by hipcc for the kernels and by a host compiler for
// the CPU accuracy tests; NOA_HD carries the qualifiers.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NOA_HD __host__ __device__
#define NOA_FM_TABLE static __constant__ const
#else
#define NOA_HD
#define NOA_FM_TABLE static const
#endif

#include "fastmath_tables.hh"

namespace noa_hip {
namespace fm {

constexpr int TABLE_DOUBLES = 3 * LOG_N + 2 * EXP_N;
// LDS alignment of the table block: ds_read2st64_b64 reads two entries 64 x 8
// bytes apart at one address with offsets in 512-byte units, so at a 512-byte
// aligned base the byte index i * 8 is the whole address (no add per call).
constexpr int TABLE_LDS_ALIGN = 512;

// Pointers to one copy of the tables (LDS in the kernels).
struct Tab {
    const double* invc;
    const double* lhi;
    const double* llo;
    const double* ehi;
    const double* elo;
};

// Layout of the flat table block: invc | logc_hi | logc_lo | exp_hi | exp_lo.
NOA_HD inline Tab tab_from_block(const double* b) {
    return Tab{b, b + LOG_N, b + 2 * LOG_N, b + 3 * LOG_N, b + 3 * LOG_N + EXP_N};
}

// Source of the flat block, element k.
NOA_HD inline double table_entry(int k) {
    if (k < LOG_N) return LOG_INVC[k];
    if (k < 2 * LOG_N) return LOG_LOGC_HI[k - LOG_N];
    if (k < 3 * LOG_N) return LOG_LOGC_LO[k - 2 * LOG_N];
    if (k < 3 * LOG_N + EXP_N) return EXP_T_HI[k - 3 * LOG_N];
    return EXP_T_LO[k - 3 * LOG_N - EXP_N];
}

// a * b + c for a compile-time constant c (fma_kc) or b (fma_kb). gfx950's
// v_fma_f64 takes one scalar operand, but the compiler shrinks an FMA whose
// addend is a constant to the two-address v_fmac_f64 and first copies the
// constant into the destination VGPR with two v_mov_b32, three VALU slots for
// one FMA. Here the constant is pinned to an SGPR pair (set by s_mov_b32 on
// the scalar unit), so each polynomial step is one VALU instruction. The
// constraint "s" makes the operand wave-uniform: pass only constants.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ double fma_kc(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
}
__device__ __forceinline__ double fma_kb(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c));
    return d;
}
// -a * b + c with constant b
__device__ __forceinline__ double fnma_kb(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, -%1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c));
    return d;
}
#else
inline double fma_kc(double a, double b, double c) { return std::fma(a, b, c); }
inline double fma_kb(double a, double b, double c) { return std::fma(a, b, c); }
inline double fnma_kb(double a, double b, double c) { return std::fma(-a, b, c); }
#endif

// 1 / x for a positive normal finite x: the hardware reciprocal (~2^-24
// relative) and two Newton steps, five VALU instructions against ~10 for a
// division. Measured on gfx950 over 2^22 x in 10^[-300, 300]: within 0.5 ulp,
// and equal to the correctly rounded 1 / x on every input. The host build
// divides.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ double recip_pos(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = std::fma(-x, r, 1.0);
    r = std::fma(r, e, r);
    e = std::fma(-x, r, 1.0);
    return std::fma(r, e, r);
}
#else
inline double recip_pos(double x) { return 1.0 / x; }
#endif

// 1 / x for a positive normal finite x, faithfully rounded (within 1 ulp,
// not always the nearest): the hardware reciprocal refined by one third-
// order step, r (1 + e + e^2) with e = 1 - x r, three VALU instructions
// against four for recip_pos's two Newton steps (the estimate's ~2^-24
// error cubed is far below an ulp; the last FMA's rounding remains). For
// reciprocals of products that only scale other quotients, where the
// correctly rounded value matters to nothing. The host build divides.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ double recip_faithful(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    const double e = std::fma(-x, r, 1.0);
    return std::fma(r, std::fma(e, e, e), r);
}
#else
inline double recip_faithful(double x) { return 1.0 / x; }
#endif

// 1 / x for a nonzero normal x of either sign whose reciprocal is normal:
// the same steps as recip_pos (the Newton steps are sign-symmetric).
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ double recip_nz(double x) { return recip_pos(x); }
#else
inline double recip_nz(double x) { return 1.0 / x; }
#endif

// a / b given rb, the correctly rounded 1 / b: q = a rb refined by one
// Markstein step, q + (a - b q) rb, with a - b q exact in the FMA. With rb
// correctly rounded and q within an ulp of a / b this rounds to the same
// double as the division (Markstein's theorem), for normal b with a normal
// reciprocal, finite a, and a quotient that is zero or normal. Three VALU
// instructions against ~10 for the IEEE division sequence, and a shared rb
// serves every quotient by the same b. The host build runs the same steps
// (its reciprocals are IEEE divisions, the values the device's equal), so
// the CPU tests of the physics exercise this path.
NOA_HD inline double div_rn(double a, double b, double rb) {
    const double q = a * rb;
    return std::fma(std::fma(-b, q, a), rb, q);
}

NOA_HD inline uint64_t as_bits(double x) {
    uint64_t u;
    std::memcpy(&u, &x, sizeof(u));
    return u;
}

NOA_HD inline double from_bits(uint64_t u) {
    double x;
    std::memcpy(&x, &u, sizeof(x));
    return x;
}

// The table split of a positive normal x (bits ix): z in [0.5, 1) with x =
// 2^(E - 1022) z, the subinterval i of z and E as a double. All the index and
// exponent work is on the high word (32-bit operations; a 64-bit shift into
// double(E) costs the compiler a two-word int64 conversion).
struct LogSplit {
    double z;
    double kd;
    int i;
};
NOA_HD inline LogSplit log_split(uint64_t ix, int kadj = 0) {
    const uint32_t hx = uint32_t(ix >> 32);
    LogSplit s;
    s.i = int((hx >> (LOG_SHIFT - 32)) & uint32_t(LOG_N - 1));
    // an arithmetic shift (the sign bit is clear): from a logical one the
    // compiler widens the conversion to a two-word int64 one
    s.kd = double((int32_t(hx) >> 20) + kadj);
#if defined(__HIP_DEVICE_COMPILE__)
    // the and-or as one v_bfi_b32 (its VOP3 form takes no literal: the mask
    // in an SGPR, the exponent word in a VGPR the compiler keeps out of the
    // loops); the compiler's own choice is a v_and and a v_or with literals.
    // Measured on 2^24 pairs: 1% faster FP64 kernel than with those two, as
    // fast as v_frexp_mant_f64 there and 0.6% faster in the mixed form
    uint32_t zh;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(zh) : "s"(0x000fffffu), "v"(hx), "v"(0x3fe00000u));
    s.z = from_bits((uint64_t(zh) << 32) | (ix & 0xffffffffull));
#else
    s.z = from_bits((uint64_t((hx & 0x000fffffu) | 0x3fe00000u) << 32) | (ix & 0xffffffffull));
#endif
    return s;
}

NOA_HD inline double log(double x, const Tab& t) {
#pragma clang fp contract(off)
    uint64_t ix = as_bits(x);
    int kadj = 0;
    // positive normal finite: 0x0010... <= ix < 0x7ff0...; the rest without libm
    if (ix - 0x0010000000000000ull >= 0x7fe0000000000000ull) {
        if (x == 0.0) return -INFINITY;
        if (!(x > 0.0)) return (x == x) ? __builtin_nan("") : x;  // negative: NaN; NaN propagates
        if (x == INFINITY) return x;
        ix = as_bits(x * 0x1p54);  // subnormal: scale into the normal range
        kadj = -54;
    }
    const LogSplit sp = log_split(ix, kadj);
    const int i = sp.i;
    const double kd = sp.kd;
    const double r = std::fma(sp.z, t.invc[i], -1.0);
    const double w = fma_kb(kd, LN2_HI, t.lhi[i]);  // exact
    double q = fma_kc(r, LOG_P4, LOG_P3);
    q = fma_kc(q, r, LOG_P2);
    q = fma_kc(q, r, LOG_P1);
    q = fma_kc(q, r, LOG_P0);
    // w exact and |r| <= 2^-8: summing the small parts first leaves two
    // roundings, <= 0.75 ulp in all (a TwoSum of w + r would buy 0.15 ulp
    // for five more additions)
    const double lo = fma_kb(kd, LN2_LO, t.llo[i]);
    return w + (r + std::fma(r * r, q, lo));
}

// log of a positive normal finite x, without the special-argument path
// (~30% cheaper at a call site). Zero, negative, subnormal or infinite x
// give an unspecified finite value: use it only where the argument is
// bounded away from those by construction.
NOA_HD inline double log_pos(double x, const Tab& t) {
#pragma clang fp contract(off)
    const LogSplit sp = log_split(as_bits(x));
    const int i = sp.i;
    const double kd = sp.kd;
    const double r = std::fma(sp.z, t.invc[i], -1.0);
    const double w = fma_kb(kd, LN2_HI, t.lhi[i]);
    double q = fma_kc(r, LOG_P4, LOG_P3);
    q = fma_kc(q, r, LOG_P2);
    q = fma_kc(q, r, LOG_P1);
    q = fma_kc(q, r, LOG_P0);
    const double lo = fma_kb(kd, LN2_LO, t.llo[i]);
    return w + (r + std::fma(r * r, q, lo));
}

NOA_HD inline double exp(double x, const Tab& t) {
#pragma clang fp contract(off)
    if (x != x) return x;
    // beyond +-750 the result is inf or 0 either way; the clamp keeps k in
    // int range, and ldexp rounds the overflow / subnormal results
    x = std::fmin(std::fmax(x, -750.0), 750.0);
    const double kd = std::rint(x * INV_LN2_N);
    const int ki = int(kd);
    double r = fnma_kb(kd, LN2_N_HI, x);
    r = fnma_kb(kd, LN2_N_LO, r);
    const int j = ki & (EXP_N - 1);
    const int m = ki >> 7;  // floor(ki / 128)
    double p = fma_kc(r, 1.0 / 120.0, 1.0 / 24.0);
    p = fma_kc(p, r, 1.0 / 6.0);
    p = std::fma(p, r, 0.5);
    p = std::fma(p, r, 1.0);
    p = p * r;
    const double th = t.ehi[j];
    const double y = th + std::fma(th, p, t.elo[j]);
    return std::ldexp(y, m);
}

// exp of a finite |x| < 2^31 ln2 / 128 (about 1.16e7), or NaN, which
// propagates, without the NaN test and the clamp; overflow and underflow
// still come out of ldexp as inf, 0 or subnormal. An infinite x gives NaN.
// k = rint(128 x / ln2) is read as a signed 32-bit integer from the low
// mantissa bits of 128 x / ln2 + 1.5 * 2^52, formed by one FMA (no
// float-to-int conversion; k may differ from the rounded product's at a tie,
// which leaves |r| within ln2 / 256 + 1 ulp, far inside the polynomial's
// range). Beyond the stated range k wraps and the result is wrong; the
// callers' arguments are node offsets and logarithms of physical ratios,
// |x| < 1e3.
NOA_HD inline double exp_fin(double x, const Tab& t) {
#pragma clang fp contract(off)
    const double shifted = std::fma(x, INV_LN2_N, 0x1.8p52);
    const int ki = int(uint32_t(as_bits(shifted)));
    const double kd = shifted - 0x1.8p52;
    double r = fnma_kb(kd, LN2_N_HI, x);
    r = fnma_kb(kd, LN2_N_LO, r);
    const int j = ki & (EXP_N - 1);
    const int m = ki >> 7;
    double p = fma_kc(r, 1.0 / 120.0, 1.0 / 24.0);
    p = fma_kc(p, r, 1.0 / 6.0);
    p = std::fma(p, r, 0.5);
    p = std::fma(p, r, 1.0);
    p = p * r;
    const double th = t.ehi[j];
    const double y = th + std::fma(th, p, t.elo[j]);
    return std::ldexp(y, m);
}

// log_pos without the r^4 term of its polynomial (the same table and
// constants otherwise): absolute error < 6.5e-16 (the dropped term is at
// most |LOG_P4| r^6 <= 6e-16 over |r| <= 2^-8), one FMA less. For
// logarithms whose absolute error is what counts (added into exponents).
NOA_HD inline double log_pos_lo(double x, const Tab& t) {
#pragma clang fp contract(off)
    const LogSplit sp = log_split(as_bits(x));
    const int i = sp.i;
    const double kd = sp.kd;
    const double r = std::fma(sp.z, t.invc[i], -1.0);
    const double w = fma_kb(kd, LN2_HI, t.lhi[i]);
    double q = fma_kc(r, LOG_P3, LOG_P2);
    q = fma_kc(q, r, LOG_P1);
    q = fma_kc(q, r, LOG_P0);
    const double lo = fma_kb(kd, LN2_LO, t.llo[i]);
    return w + (r + std::fma(r * r, q, lo));
}

// exp_fin with a polynomial one degree shorter: e^r - 1 = r (1 + r Q(r)),
// Q the degree-2 Chebyshev economization of (e^r - 1 - r) / r^2 over the
// reduced range (truncation < 1.4 ulp; measured within 3 ulp in all against
// 0.6 for exp_fin), one FMA less. Same argument range as exp_fin. For factors whose
// relative error passes through to the result unamplified.
NOA_HD inline double exp_fin4(double x, const Tab& t) {
#pragma clang fp contract(off)
    const double shifted = std::fma(x, INV_LN2_N, 0x1.8p52);
    const int ki = int(uint32_t(as_bits(shifted)));
    const double kd = shifted - 0x1.8p52;
    double r = fnma_kb(kd, LN2_N_HI, x);
    r = fnma_kb(kd, LN2_N_LO, r);
    const int j = ki & (EXP_N - 1);
    const int m = ki >> 7;
    double p = fma_kc(r, EXP4_P2, EXP4_P1);
    p = fma_kc(p, r, EXP4_P0);
    p = std::fma(p, r, 1.0);
    p = p * r;
    const double th = t.ehi[j];
    const double y = th + std::fma(th, p, t.elo[j]);
    return std::ldexp(y, m);
}

}  // namespace fm
}  // namespace noa_hip
