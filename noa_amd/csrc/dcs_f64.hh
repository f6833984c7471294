// FP64 device physics of the four DCS processes.
//
// FMA contraction is off in every function body and the closed forms keep
// the reference's evaluation order. The node loops of the two integrals are
// rewritten algebraically to spend fewer divisions, square roots and
// logarithms (the same mathematics with a few ulps of different rounding,
// far inside the 1e-10 parity bound):
//   pair production  the two screening logarithms of L_e become one,
//                    0.5 log(A^2 Z^-2/3 q^2 rho21^2 x_e / ((1 + c_Le x_e)
//                    (q rho21 + c_L x_e)^2)), with no square root; c_L / rho21
//                    is folded into the log arguments; (rho21 - beta) / (1 +
//                    xi) is shared by B_e and B_mu; / r^2 is a product with
//                    1 / r^2; each node takes a correctly rounded 1 / xi (the
//                    Newton reciprocal, the reference's quotient) and
//                    two reciprocals of products (the two Y denominators with
//                    1 + xi; the two screening denominators) where the
//                    reference divides five more times;
//   photonuclear     W^2 - M^2 + Q^2 = 2 M q, so the ALLM denominators of x_P
//                    and x_R are per-pair constants; one per-pair reciprocal
//                    gives 1 / q, 1 / E^2 and those two; in a node, Whitlow's
//                    R and theta are carried as fractions, so that 1 / (1 + R)
//                    is D_R D_theta / S, and one reciprocal of Q^2 (Q^2 + m0^2)
//                    g_P g_A S replaces the reference's eight divisions; log10
//                    as log / ln 10; its sums of products are FMAs where
//                    they save an instruction.
// Inside the node loops, log_pos / exp_fin (no special-argument paths) take
// the arguments that are positive normal / finite by construction for total
// energies up to 1e30 GeV: 1 + 1/xi, 3 + xi, the screening arguments, Q^2's
// exponent, (Q^2 + Q_0^2) / Lambda^2 >= 8, 1 - x >= m_pi / q, x_P, x_R and
// the ALLM exponents. NaN inputs still give NaN results through the
// arithmetic around them. Above 1e30 GeV, for infinite or NaN K, and for an
// element outside the fast_ok range, the integrals run with the full log /
// exp and IEEE divisions (the same code, templated), so the reference's
// overflow results are kept.
// The quotients of the closed forms, the exits and the integrals' set-up
// (a / b in the reference) are fm::div_rn(a, b, 1 / b): a correctly rounded
// reciprocal refined by one Markstein step, which rounds as the division
// does for the normal operands that `ordinary` (K, q) and the element's
// fast_ok flag guarantee, at about a third of the IEEE sequence's cost; other
// inputs take the same code with the divisions (the Safe instantiations).
// Otherwise the device log/exp (the table-driven ones of fastmath.hh, ~0.5
// ulp) against glibc is the only difference from the reference's CPU path.
// Launch invariants come precomputed in DcsConsts.
//   bremsstrahlung   src/noa/pms/physics.hh:114-153
//   pair production  src/noa/pms/dcs.hh:144-258 (8-node rule numerics.hh:110-129)
//   photonuclear     src/noa/pms/dcs.hh:261-405 (9-node rule numerics.hh:131-152)
//   ionisation       src/noa/pms/dcs.hh:408-443
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NOA_DEV __device__
#define NOA_DEV_INLINE __device__ __forceinline__
#else  // host build of the same arithmetic: the CPU-tensor path (dcs_host.cc) and the CPU tests
#include <cmath>
#define NOA_DEV
#define NOA_DEV_INLINE inline
#endif

#include "dcs_consts.hh"
#include "fastmath.hh"

// FMA contraction is switched off inside each function body (block scope),
// so including this header does not change the rest of a translation unit.

namespace noa_hip {

// Gauss-Legendre node tables; the loop index is wave-uniform, so these are
// scalar loads. Values: numerics.hh:116-121 (8 nodes on [0,1], truncated
// literals kept as specified) and numerics.hh:137-144 (9 nodes on [-1,1]).
constexpr double INV_LN10 = 0.43429448190325182765;  // 1 / ln 10

NOA_FM_TABLE double GL8_X[8] = {0.01985507, 0.10166676, 0.2372338, 0.40828268,
                                0.59171732, 0.7627662, 0.89833324, 0.98014493};
NOA_FM_TABLE double GL8_W[8] = {0.05061427, 0.11119052, 0.15685332, 0.18134189,
                                0.18134189, 0.15685332, 0.11119052, 0.05061427};
NOA_FM_TABLE double GL9_X[9] = {0.0000000000000000, -0.8360311073266358, 0.8360311073266358,
                                -0.9681602395076261, 0.9681602395076261, -0.3242534234038089,
                                0.3242534234038089, -0.6133714327005904, 0.6133714327005904};
NOA_FM_TABLE double GL9_W[9] = {0.3302393550012598, 0.1806481606948574, 0.1806481606948574,
                                0.0812743883615744, 0.0812743883615744, 0.3123470770400029,
                                0.3123470770400029, 0.2606106964029354, 0.2606106964029354};

// NU: unroll factor of the integrals' node loops (1 or 2), a choice of the
// batch autotuner; F64 is the plain form.
template <int NU>
struct F64N {
    using out_t = double;
    static constexpr bool has_redo = false;

    // Quotients of the closed forms. On the fast path a / b is div_rn(a, b,
    // 1 / b) with 1 / b from recip_nz, both rounding as the IEEE division
    // does, so the results (and every exit decision taken on them) are the
    // reference's; the Safe instantiation divides. `ordinary` selects the
    // fast path per lane: 1e-20 <= K <= 1e30 and 1e-30 <= q <= 1e30 GeV
    // (NaN fails), for an element whose constants allow it. Under those
    // bounds every divisor and quotient below is normal (or an exact zero
    // numerator), which div_rn needs.
    template <bool Safe>
    static NOA_DEV_INLINE double rcp(double b) {
        if constexpr (Safe) return 1. / b;
        else return fm::recip_nz(b);
    }
    template <bool Safe>
    static NOA_DEV_INLINE double dv(double a, double b, double rb) {
        if constexpr (Safe) return a / b;
        else return fm::div_rn(a, b, rb);
    }
    static NOA_DEV_INLINE bool ordinary(double K, double q, const DcsConsts& c) {
        return c.fast_ok && (K >= 1e-20) && (K <= FAST_E_MAX) && (q >= 1e-30) && (q <= FAST_E_MAX);
    }
    static constexpr double R_ME = 1. / ELECTRON_MASS;  // correctly rounded at compile time

    // physics.hh:114-153
    static NOA_DEV_INLINE double brems(double K, double q, const DcsConsts& c, const fm::Tab& t) {
#pragma clang fp contract(off)
        // 1 - nu >= 2^-20 keeps delta's divisor normal on the fast path
        if (ordinary(K, q, c) && q <= (K + c.mass) * (1. - 0x1p-20)) return brems_impl<false>(K, q, c, t);
        return brems_impl<true>(K, q, c, t);
    }

    template <bool Safe>
    static NOA_DEV_INLINE double brems_impl(double K, double q, const DcsConsts& c, const fm::Tab& t) {
#pragma clang fp contract(off)
        const double me = ELECTRON_MASS;
        const double sqrte = SQRTE_BREMS;
        const double E = K + c.mass;
        const double rE = Safe ? 0. : rcp<Safe>(E);
        const double delta_factor = dv<Safe>(c.b_half_m2, E, rE);
        const double meE = me * E;
        const double t1 = dv<Safe>(c.b_half_m2, meE, Safe ? 0. : rcp<Safe>(meE));
        const double d1 = 1. + t1;
        const double qe_max = dv<Safe>(E, d1, Safe ? 0. : rcp<Safe>(d1));
        const double nu = dv<Safe>(q, E, rE);
        const double omn = 1. - nu;
        const double delta = dv<Safe>(delta_factor * nu, omn, Safe ? 0. : rcp<Safe>(omn));
        const double Nn = c.b_BZ_n * (c.mass + delta * c.b_Dn_sqrte_m2);
        const double Dn = c.b_D_n * (me + delta * sqrte * c.b_BZ_n);
        // fast path: positive normal arguments (N_n, D_n, D_e > 0), so log_pos
        double Phi_n = nlog<Safe>(dv<Safe>(Nn, Dn, Safe ? 0. : rcp<Safe>(Dn)), t);
        if (Phi_n < 0.) Phi_n = 0.;
        double Phi_e = 0.;
        if (q < qe_max) {
            const double De = (1. + delta * c.b_phie_factor) * (me + delta * sqrte * c.b_BZ_e);
            Phi_e = nlog<Safe>(dv<Safe>(c.b_BZe_mass, De, Safe ? 0. : rcp<Safe>(De)), t);
            if (Phi_e < 0.) Phi_e = 0.;
        }
        const double dcs = c.b_dcs_factor * (c.Zd * Phi_n + Phi_e) * (4. / 3. * (rcp<Safe>(nu) - 1.) + nu);
        return (dcs < 0.) ? 0. : dv<Safe>(dcs * 1E+03 * AVOGADRO_NUMBER, c.A, c.r_A);
    }

    // dcs.hh:408-443
    static NOA_DEV_INLINE double ion(double K, double q, const DcsConsts& c, const fm::Tab& t) {
        if (ordinary(K, q, c)) return ion_impl<false>(K, q, c, t);
        return ion_impl<true>(K, q, c, t);
    }

    // The same for the mixed-precision kernel's FP64 recomputation, with the
    // full log for 4 E (E - q) / m^2 (its register allocation is the tighter
    // one; the value is the same)
    static NOA_DEV_INLINE double ion_redo(double K, double q, const DcsConsts& c, const fm::Tab& t) {
        if (ordinary(K, q, c)) return ion_impl<false, false>(K, q, c, t);
        return ion_impl<true>(K, q, c, t);
    }

    template <bool Safe, bool Lean = !Safe>
    static NOA_DEV_INLINE double ion_impl(double K, double q, const DcsConsts& c, const fm::Tab& t) {
#pragma clang fp contract(off)
        const double me = ELECTRON_MASS;
        const double P2 = K * (K + c.i_2m);
        const double E = K + c.mass;
        const double Dw = c.i_m2 + me * (me + 2. * E);
        const double Wmax = dv<Safe>(2. * me * P2, Dw, Safe ? 0. : rcp<Safe>(Dw));
        if ((Wmax < X_FRACTION * K) || (q > Wmax)) return 0.;
        if (q <= c.i_Wmin) return 0.;
        const double rP2 = Safe ? 0. : rcp<Safe>(P2);
        const double a0 = Safe ? 0.5 / P2 : 0.5 * rP2;  // a power-of-two multiple: the same rounding
        const double a1 = -rcp<Safe>(Wmax);
        const double a2 = dv<Safe>(E * E, P2, rP2);
        const double rq = rcp<Safe>(q);
        const double cs = dv<Safe>(1.535336E-05 * E * c.Zd, c.A, c.r_A) * (a0 + rq * (a1 + dv<Safe>(a2, q, rq)));
        double Delta = 0.;
        if (K >= c.i_Kthr) {
            // 1 + 2q/m_e >= 1: log_pos on the fast path. E - q can round to 0
            // (E = K + m == K for K >> m, q = Wmax = K), where the reference's
            // log gives -inf; otherwise E - q >= ulp(E) keeps 4 E (E - q) / m^2
            // positive normal, so the fast path needs no other special case
            const double L1 = nlog<Safe>(1. + dv<Safe>(2. * q, me, R_ME), t);
            const double a4 = dv<Safe>(4. * E * (E - q), c.i_m2, c.i_r_m2);
            const double L4 = Lean ? ((a4 > 0.) ? fm::log_pos(a4, t) : -INFINITY) : fm::log(a4, t);
            Delta = 1.16141E-03 * L1 * (L4 - L1);
        }
        return cs * (1. + Delta);
    }

    // dcs.hh:151-156 and 168-175: the exits before the integral.
    static NOA_DEV_INLINE bool pair_active(double K, double q, const DcsConsts& c) {
        if (ordinary(K, q, c)) return pair_active_impl<false>(K, q, c);
        return pair_active_impl<true>(K, q, c);
    }

    // The same, also returning the lower end of the integral, argmin (the
    // value pair_impl computes for it: both quotient forms round as the
    // division for an active pair's normal operands), for pair_at.
    static NOA_DEV_INLINE bool pair_active(double K, double q, const DcsConsts& c, double& argmin) {
        if (ordinary(K, q, c)) return pair_active_impl<false>(K, q, c, &argmin);
        return pair_active_impl<true>(K, q, c, &argmin);
    }

    template <bool Safe>
    static NOA_DEV_INLINE bool pair_active_impl(double K, double q, const DcsConsts& c, double* argmin_out = nullptr) {
#pragma clang fp contract(off)
        if (q <= 4. * ELECTRON_MASS) return false;
        if (q >= K + c.p_cut) return false;
        const double gamma = 1. + dv<Safe>(K, c.mass, c.r_mass);
        const double x0 = dv<Safe>(4. * ELECTRON_MASS, q, Safe ? 0. : rcp<Safe>(q));
        const double gg = gamma * (gamma - dv<Safe>(q, c.mass, c.r_mass));
        const double x1 = dv<Safe>(6., gg, Safe ? 0. : rcp<Safe>(gg));
        const double Na = x0 + 2. * (1. - x0) * x1;
        double Da = 1. + (1. - x1) * sqrt(1. - x0);
        if constexpr (!Safe) {
            // |Da| < 2^-900 (also a signed zero) makes |argmin| > 2^800 or
            // infinite in the reference: an exit either way, and a positive
            // 2^-900 gives the same exit without a zero divisor
            if (fabs(Da) < 0x1p-900) Da = 0x1p-900;
        }
        const double argmin = dv<Safe>(Na, Da, Safe ? 0. : rcp<Safe>(Da));
        if (argmin_out) *argmin_out = argmin;
        return !((argmin >= 1.) || (argmin <= 0.));
    }

    // Node-loop log / exp: the full versions on the safe path, the ones
    // without special-argument handling otherwise.
    template <bool Safe>
    static NOA_DEV_INLINE double nlog(double x, const fm::Tab& t) {
        return Safe ? fm::log(x, t) : fm::log_pos(x, t);
    }
    template <bool Safe>
    static NOA_DEV_INLINE double nexp(double x, const fm::Tab& t) {
        return Safe ? fm::exp(x, t) : fm::exp_fin(x, t);
    }

    // Total energy up to which every node-loop argument of the two integrals
    // stays positive normal / finite (bounds: beta >= 1e-66, the screening
    // denominators <= 1e140, 1 - x >= 1e-31).
    static constexpr double FAST_E_MAX = 1e30;

    // dcs.hh:144-258 for a pair that passed pair_active.
    static NOA_DEV double pair(double K, double q, const DcsConsts& c, const fm::Tab& t) {
        if (c.fast_ok && K + c.mass <= FAST_E_MAX) return pair_impl<false>(K, q, c, t);
        return pair_impl<true>(K, q, c, t);  // also infinite and NaN K
    }

    // The same from the argmin pair_active returned (no second evaluation of
    // the exit's quotients and square root); the same result.
    static NOA_DEV double pair_at(double K, double q, double argmin, const DcsConsts& c, const fm::Tab& t) {
        if (c.fast_ok && K + c.mass <= FAST_E_MAX) return pair_impl<false, true, true>(K, q, c, t, argmin);
        return pair_impl<true, false, true>(K, q, c, t, argmin);
    }

    // The same with the set-up quotients as IEEE divisions (Rn = false): the
    // FP64 recomputation inside the mixed-precision kernel, where the
    // div_rn path's extra constants raise that kernel's scalar register
    // spills by more than the divisions cost on ~1% of its pairs.
    static NOA_DEV double pair_ieee(double K, double q, const DcsConsts& c, const fm::Tab& t) {
        if (K + c.mass <= FAST_E_MAX) return pair_impl<false, false>(K, q, c, t);
        return pair_impl<true>(K, q, c, t);
    }

    template <bool Safe, bool Rn = !Safe, bool HaveArg = false>
    static NOA_DEV double pair_impl(double K, double q, const DcsConsts& c, const fm::Tab& t, double argmin_in = 0.) {
#pragma clang fp contract(off)
        const double me = ELECTRON_MASS;
        // the set-up quotients as in pair_active (a pair that passed it has
        // 4 m_e < q < K + p_cut, p_cut < 0, so K, q > 2e-3 GeV, E - q > 0.76 m
        // and argmin in (0, 1): every divisor below is positive normal)
        const double E0 = K + c.mass;
        const double nu = dv<!Rn>(q, E0, !Rn ? 0. : rcp<!Rn>(E0));
        const double omn = 1. - nu;
        const double beta = dv<!Rn>(0.5 * nu * nu, omn, !Rn ? 0. : rcp<!Rn>(omn));
        const double xi_factor = c.p_hr2 * beta;
        const double gamma = 1. + dv<!Rn>(K, c.mass, c.r_mass);
        double argmin = argmin_in;
        if constexpr (!HaveArg) {
            const double x0 = dv<!Rn>(4. * me, q, !Rn ? 0. : rcp<!Rn>(q));
            const double gg = gamma * (gamma - dv<!Rn>(q, c.mass, c.r_mass));
            const double x1 = dv<!Rn>(6., gg, !Rn ? 0. : rcp<!Rn>(gg));
            const double Da = 1. + (1. - x1) * sqrt(1. - x0);
            argmin = dv<!Rn>(x0 + 2. * (1. - x0) * x1, Da, !Rn ? 0. : rcp<!Rn>(Da));
        }
        const double tmin = nlog<Safe>(argmin, t);  // argmin in (0, 1), >= x0 / 2 > 1e-33

        // per-pair coefficients of the node polynomials in rho^2 (explicit
        // FMAs: with contraction off the compiler keeps every product and
        // sum of the reference's written order)
        const double c1b = 1. + beta, c15b = 1. + 1.5 * beta, c2b = 1. + 2. * beta;
        const double Ne0 = 5. + 4. * beta, Ne1 = 4. * beta - 1.;  // N_e = 5 - rho2 + 4 beta (1 + rho2)
        const double Nm0 = 4. + 3. * beta, Nm1 = 1. + 3. * beta;  // N_mu = 4 + rho2 + 3 beta (1 + rho2)
        const double De0 = -4. * beta, De1 = 2. * beta - 1., De2 = 2. * Nm1;  // D_e = De2 L - rho2 - 2 beta (2 - rho2)
        const double Dm2 = 1.5 + 2. * beta;

        double I = 0.;
#pragma unroll NU
        for (int j = 0; j < 8; j++) {
            const double tj = GL8_X[j];
            const double eps = nexp<Safe>(tj * tmin, t);
            const double rho = 1. - eps;
            const double rho2 = rho * rho;
            const double opr2 = 1. + rho2;
            const double rho21 = eps * (2. - eps);
            const double xi = xi_factor * rho21;
            const double xi1 = 1. + xi;
            // correctly rounded, as in the reference: near xi = 1e3, B_e's
            // log(1 + 1/xi) - ... cancels ~3000-fold, so one ulp of 1/xi that
            // changes the rounding of 1 + 1/xi moves B_e by ~1e-10. recip_pos
            // ends in a Newton step from an approximation within 2^-46, so it
            // can round differently from 1 / xi only where 1 / xi lies within
            // ~2^-92 relative of a rounding midpoint (~2^-39 of inputs); xi >=
            // 1e-66 here, so xi is positive normal
            const double xi_i = Safe ? 1. / xi : fm::recip_pos(xi);
            // log(3 + 1/xi) and log(3 + xi) enter D_e and D_mu, the screening
            // logs below Phi_e and Phi_mu, all through their absolute error: on
            // the fast path log_pos_lo (< 6.5e-16 absolute). The logs of B_e and
            // B_mu, whose terms cancel, keep log_pos
            const double L3e = Safe ? nlog<Safe>(3. + xi_i, t) : fm::log_pos_lo(3. + xi_i, t);
            const double L3m = Safe ? nlog<Safe>(3. + xi, t) : fm::log_pos_lo(3. + xi, t);
            double De, Dm, Ne, Nm;
            if constexpr (Safe) {  // the reference's order (its overflow and NaN results)
                De = 2. * (1. + 3. * beta) * L3e - rho2 - 2. * beta * (2. - rho2);
                Dm = (1. + rho2) * (1.5 + 2. * beta) * L3m + 1. - 1.5 * rho2;
                Ne = 5. - rho2 + 4. * beta * (1. + rho2);
                Nm = 4. + rho2 + 3. * beta * (1. + rho2);
            } else {
                De = fma(De2, L3e, fma(rho2, De1, De0));
                Dm = fma(opr2 * Dm2, L3m, fma(-1.5, rho2, 1.));
                Ne = fma(rho2, Ne1, Ne0);
                Nm = fma(rho2, Nm1, Nm0);
            }
            // dr = (rho21 - beta) / (1 + xi), shared by B_e and B_mu; x_e = (1 + xi)
            // (1 + Y_e), x_mu = (1 + xi) (1 + Y_mu) with Y_e = N_e / D_e, Y_mu = N_mu / D_mu
            double dr, xe, xmu;
            if constexpr (Safe) {
                dr = (rho21 - beta) / xi1;
                xe = xi1 * (1. + Ne / De);
                xmu = xi1 * (1. + Nm / Dm);
            } else {  // one reciprocal for the three (D_e, D_mu > 1)
                const double DD = De * Dm;
                const double r_y = Rn ? fm::recip_faithful(DD * xi1) : fm::recip_pos(DD * xi1);
                dr = (rho21 - beta) * (DD * r_y);
                if constexpr (Rn) {
                    const double u = xi1 * (xi1 * r_y);  // (1 + xi) / (D_e D_mu)
                    xe = fma(Ne * Dm, u, xi1);
                    xmu = fma(Nm * De, u, xi1);
                } else {  // (the mixed form's recomputation: fewer registers live)
                    xe = xi1 * (1. + Ne * ((Dm * xi1) * r_y));
                    xmu = xi1 * (1. + Nm * ((De * xi1) * r_y));
                }
            }
            // B_e and B_mu keep the reference's order: near xi = 1e3 and 1e-3
            // their terms cancel, and any reordering moves the result by up to
            // ~4e-11 there
            double Be;
            if (xi >= 1E+03)
                Be = 0.5 * xi_i * ((3. - rho2) + 2. * beta * (1. + rho2));
            else
                Be = ((2. + rho2) * c1b + xi * (3. + rho2)) * nlog<Safe>(1. + xi_i, t) + dr - 3. - rho2;
            double Bmu;
            if (xi <= 1E-03)
                Bmu = 0.5 * xi * (5. - rho2 + beta * (3. + rho2));
            else
                Bmu = (opr2 * c15b - xi_i * c2b * rho21) * nlog<Safe>(1. + xi, t) + xi * dr + c2b * rho21;
            // q + (c_L / rho21) x = (q rho21 + c_L x) / rho21
            const double qr = q * rho21;
            if constexpr (Safe) {
                const double ge = qr + c.p_cL * xe;
                const double Le = fm::log(c.p_AZ13 * qr * sqrt(xe / ((1. + c.p_cLe * xe) * (ge * ge))), t);
                const double Lmu = fm::log(c.p_rAZ13 * qr / (c.p_15Z13 * (qr + c.p_cL * xmu)), t);
                double Phi_e = Be * Le;
                double Phi_mu = Bmu * Lmu;
                if (Phi_e < 0.) Phi_e = 0.;
                if (Phi_mu < 0.) Phi_mu = 0.;
                const double f = -(Phi_e + Phi_mu * c.p_inv_r2) * (1. - rho) * tmin;
                I += f * GL8_W[j];
            } else {
                // one reciprocal for the two screening denominators, no square
                // root; L_e2 = 2 L_e, so Phi_e2 = 2 Phi_e exactly, and the node
                // sums 2 (Phi_e + Phi_mu / r^2) (1 - rho) w_j, scaled by
                // -tmin / 2 after the loop (in the mixed form's recomputation,
                // which keeps fewer registers live, each node's sum is scaled
                // by -tmin and Phi_e is formed directly). 1 - rho as the reference rounds it
                // (not eps: they differ by ~1e-16 absolute, which is ~1e-11
                // relative where eps ~ 1e-5)
                const double ge = fma(c.p_cL, xe, qr);
                const double Dle = fma(c.p_cLe, xe, 1.) * (ge * ge);
                const double Dlm = c.p_15Z13 * fma(c.p_cL, xmu, qr);
                const double r_l = Rn ? fm::recip_faithful(Dle * Dlm) : fm::recip_pos(Dle * Dlm);
                const double Le2 = fm::log_pos_lo(c.p_AZ13_sq * (qr * qr) * xe * (Dlm * r_l), t);
                const double Lmu = fm::log_pos_lo(c.p_rAZ13 * qr * (Dle * r_l), t);
                // one v_max each; a NaN here comes from a NaN q, which the
                // (E - q) / q factor of the result carries anyway
                const double Phi_mu = fmax(Bmu * Lmu, 0.);
                if constexpr (Rn) {
                    const double Phi_e2 = fmax(Be * Le2, 0.);
                    I = fma(fma(Phi_mu, 2. * c.p_inv_r2, Phi_e2), (1. - rho) * GL8_W[j], I);
                } else {
                    const double Phi_e = fmax(Be * (0.5 * Le2), 0.);
                    I = fma(fma(Phi_mu, c.p_inv_r2, Phi_e), (1. - rho) * (-tmin * GL8_W[j]), I);
                }
            }
        }
        if constexpr (Rn) I *= -0.5 * tmin;

        double zeta;
        if (gamma <= 35.) {
            zeta = 0.;
        } else {
            const double z1 = 1. + c.p_g1 * gamma * c.p_Z13 * c.p_Z13;
            zeta = 0.073 * nlog<Safe>(dv<!Rn>(gamma, z1, !Rn ? 0. : rcp<!Rn>(z1)), t) - 0.26;
            if (zeta <= 0.) {
                zeta = 0.;
            } else {
                const double z2 = 1. + c.p_g2 * gamma * c.p_Z13;
                // the divisor 0.058 log(.) - 0.14 can be zero: a plain division
                zeta /= 0.058 * nlog<Safe>(dv<!Rn>(gamma, z2, !Rn ? 0. : rcp<!Rn>(z2)), t) - 0.14;
            }
        }
        const double E = K + c.mass;
        const double qE = q * E;
        const double dcs = dv<!Rn>(c.p_dcs_c * (c.Zd + zeta) * (E - q) * I, qE, !Rn ? 0. : rcp<!Rn>(qE));
        return (dcs < 0.) ? 0. : dv<!Rn>(dcs * 1E+03 * AVOGADRO_NUMBER * (c.mass + K), c.A, c.r_A);
    }

    // dcs.hh:357-381: the exits before the integral.
    static NOA_DEV_INLINE bool photo_active(double K, double q, const DcsConsts& c) {
        if (ordinary(K, q, c)) return photo_active_impl<false>(K, q, c);
        return photo_active_impl<true>(K, q, c);
    }

    template <bool Safe>
    static NOA_DEV_INLINE bool photo_active_impl(double K, double q, const DcsConsts& c) {
#pragma clang fp contract(off)
        if ((q < 1.) || (q < 2E-03 * K)) return false;
        const double E = K + c.mass;
        if ((q >= (E - c.mass)) || (q <= c.ph_thr_pi)) return false;
        // past the exits 1 <= q < K, so y and 1 - y >= m / E are normal
        const double y = dv<Safe>(q, E, Safe ? 0. : rcp<Safe>(E));
        const double omy = 1. - y;
        const double Q2min = dv<Safe>(c.ph_m2 * y * y, omy, Safe ? 0. : rcp<Safe>(omy));
        const double Q2max = 2.0 * PHOTO_M * (q - PHOTO_MPI) - PHOTO_MPI * PHOTO_MPI;
        return !((Q2max < Q2min) | (Q2min < 0.));
    }

    // dcs.hh:261-405 for a pair that passed photo_active.
    static NOA_DEV double photo(double K, double q, const DcsConsts& c, const fm::Tab& t) {
        if (c.fast_ok && K + c.mass <= FAST_E_MAX) return photo_impl<false>(K, q, c, t);
        return photo_impl<true>(K, q, c, t);
    }

    static NOA_DEV double photo_ieee(double K, double q, const DcsConsts& c, const fm::Tab& t) {
        if (K + c.mass <= FAST_E_MAX) return photo_impl<false, false>(K, q, c, t);
        return photo_impl<true>(K, q, c, t);
    }

    template <bool Safe, bool Rn = !Safe>
    static NOA_DEV double photo_impl(double K, double q, const DcsConsts& c, const fm::Tab& t) {
#pragma clang fp contract(off)
        const double M = PHOTO_M;
        const double mpi = PHOTO_MPI;
        const double E = K + c.mass;
        const double y = dv<!Rn>(q, E, !Rn ? 0. : rcp<!Rn>(E));
        const double omy = 1. - y;
        const double Q2min = dv<!Rn>(c.ph_m2 * y * y, omy, !Rn ? 0. : rcp<!Rn>(omy));
        const double Q2max = 2.0 * M * (q - mpi) - mpi * mpi;
        // 0 < Q2min < Q2max (photo_active), Q2min >= m^2 y^2 > 1e-72 below FAST_E_MAX
        const double pQ2min = nlog<Safe>(Q2min, t);
        const double pQ2max = nlog<Safe>(Q2max, t);
        const double dpQ2 = pQ2max - pQ2min;
        const double pQ2c = 0.5 * (pQ2max + pQ2min);
        const double hd = 0.5 * dpQ2;
        const double EE = E * E;
        const double twoMq = 2. * M * q;
        // x_P = (Q^2 + m_P^2) / (Q^2 + m_P^2 + W^2 - M^2) with W^2 - M^2 = 2 M q - Q^2
        // (dcs.hh:290-293): per-pair denominators
        const double dP = twoMq + 49.457, dR = twoMq + 0.15052;
        double inv_q, inv_EE, inv_dP, inv_dR;
        if constexpr (Safe) {
            inv_q = 1. / q;
            inv_EE = 1. / EE;
            inv_dP = 1. / dP;
            inv_dR = 1. / dR;
        } else {  // the four over one reciprocal
            const double qE = q * EE, PR = dP * dR;
            const double r_pair = fm::recip_pos(qE * PR);
            inv_q = EE * (PR * r_pair);
            inv_EE = q * (PR * r_pair);
            inv_dP = dR * (qE * r_pair);
            inv_dR = dP * (qE * r_pair);
        }
        const double inv_2Mq = inv_q * (0.5 / M);
        const double l_inv_2Mq = Rn ? fm::log_pos(inv_2Mq, t) : 0.;  // 1 / (2 M q) is positive normal here
        const double inv_l2 = 1. / ALLM_LAMBDA2;
        // per-pair terms of the node expressions' fused forms (fast path):
        // x_P = Q^2 / d_P + 49.457 / d_P, the same for x_R, y^2 + Q^2 / E^2,
        // 1 - y and 1 / (4 E^2)
        const double cxP = 49.457 * inv_dP, cxR = 0.15052 * inv_dR;
        const double yy = y * y, q25 = 0.25 * inv_EE;

        // the node's exponentials: in the FP64 kernel the shorter exp_fin4
        // (within 3 ulp: each feeds a factor of the integrand, relative error
        // passed on unamplified), elsewhere the node-loop exp
        const auto pexp = [](double v, const fm::Tab& tb) {
            if constexpr (Rn)
                return fm::exp_fin4(v, tb);
            else
                return nexp<Safe>(v, tb);
        };
        const auto plog = [](double v, const fm::Tab& tb) {
            if constexpr (Rn)
                return fm::log_pos_lo(v, tb);
            else
                return nlog<Safe>(v, tb);
        };
        double ds = 0.;
#pragma unroll NU
        for (int j = 0; j < 9; j++) {
            // the fast path fuses products into their sums (FMAs with at most
            // one literal operand: gfx950's VOP3 form takes one scalar and no
            // literal, so a second literal would cost two moves)
            const double lQ2 = Safe ? pQ2c + hd * GL9_X[j] : fma(hd, GL9_X[j], pQ2c);  // log Q^2 of the node
            const double Q2 = pexp(lQ2, t);
            // dcs_photonuclear_d2 (dcs.hh:335-355)
            const double x = Q2 * inv_2Mq;
            // dcs_photonuclear_f2_allm (dcs.hh:261-307)
            const double tl = plog(plog((Q2 + ALLM_Q02) * inv_l2, t) * c.ph_inv_L0, t);
            const double xP = Safe ? (Q2 + 49.457) * inv_dP : fma(Q2, inv_dP, cxP);
            const double xR = Safe ? (Q2 + 0.15052) * inv_dR : fma(Q2, inv_dR, cxR);
            // t -> 0 (Q^2 -> 0) gives log t = -inf and t^c = 0; clamped at -745
            // every t^c still rounds to 0 against the constants it is added to.
            // The FP64 kernel's fast path clamps t to the smallest normal instead (zero,
            // negative and NaN t included): log t >= -708.4 there, where every
            // t^c rounds to 0 against those constants too, so the results are
            // the same and the log needs no special-argument path
            const double lnt = Rn ? fm::log_pos_lo(fmax(tl, 0x1p-1022), t) : fmax(fm::log(tl, t), -745.0);
            const double gP = 1.0 + pexp(2.1979 * lnt, t), gA = 1.0 + pexp(1.1709 * lnt, t);
            const double bP = 0.36292 + 1.8917 * pexp(1.8439 * lnt, t);
            const double cR = 0.80107 + 0.97307 * pexp(3.4942 * lnt, t);
            const double aR = 0.58400 + 0.37888 * pexp(2.6063 * lnt, t);
            const double bR = 0.01147 + 3.7582 * pexp(0.49338 * lnt, t);
            // dcs_photonuclear_r_whitlow (dcs.hh:322-332)
            double q2 = Q2;
            if constexpr (Safe) {
                if (Q2 < 0.3) q2 = 0.3;
            } else {  // a NaN Q^2 (NaN q) still reaches the result through Q^2 itself
                q2 = fmax(Q2, 0.3);
            }
            // log(q2 / 0.04); in the FP64 kernel from the node's log Q^2 (the
            // exponent of Q^2 itself, so one log less per node): ln 25 + log Q^2
            // above the clamp, ln 7.5 at it
            const double lg = Rn ? ((Q2 >= 0.3) ? lQ2 + 3.2188758248682006 : 2.0149030205422647)
                                 : nlog<Safe>(q2 * 25.0, t);
            const double den = Safe ? 0.09 + q2 * q2 : fma(q2, q2, 0.09);
            const double Dth = (1.0 + q2) * (Safe ? 0.015625 + x * x : fma(x, x, 0.015625));
            const double Qm = Q2 + 0.31985;
            const double gPA = gP * gA;
            double inv_Q2, inv_Qm, rG, inv_1R;  // 1 / Q^2, 1 / (Q^2 + m0^2), 1 / (g_P g_A), 1 / (1 + R)
            if constexpr (Safe) {
                const double theta = 1. + (12.0 * 0.015625) * q2 / Dth;
                const double R = 0.635 / lg * theta + 0.5747 / q2 - 0.3534 / den;
                inv_1R = 1. / (1. + R);
                inv_Q2 = 1. / Q2;
                inv_Qm = 1. / Qm;
                rG = 1. / gPA;
            } else {
                // theta = (D_th + 12 c q2) / D_th and R = N_R / D_R as fractions,
                // so 1 / (1 + R) = D_R D_th / S; the four over one reciprocal
                const double DR = lg * q2 * den;
                const double S = fma(Dth, fma(lg, fma(0.5747, den, -0.3534 * q2), DR),
                                     0.635 * q2 * den * fma(12.0 * 0.015625, q2, Dth));
                const double QQm = Q2 * Qm, gS = gPA * S;
                const double r = Rn ? fm::recip_faithful(QQm * gS) : fm::recip_pos(QQm * gS);
                const double rq = gS * r, rg = QQm * r;
                inv_Q2 = Qm * rq;
                inv_Qm = Q2 * rq;
                rG = S * rg;
                inv_1R = DR * Dth * (gPA * rg);
            }
            const double cP = 0.28067 + (0.28067 - 0.22291) * (Safe ? gA * rG - 1.0 : fma(gA, rG, -1.0));
            const double aP = -0.0808 + (-0.0808 - -0.44812) * (Safe ? gP * rG - 1.0 : fma(gP, rG, -1.0));
            const double l1x = plog(1. - x, t);
            const double lxP = plog(xP, t), lxR = plog(xR, t);
            const double F2P = cP * pexp(Safe ? aP * lxP + bP * l1x : fma(aP, lxP, bP * l1x), t);
            const double F2R = cR * pexp(Safe ? aR * lxR + bR * l1x : fma(aR, lxR, bR * l1x), t);
            const double F2p = Q2 * inv_Qm * (F2P + F2R);
            // dcs_photonuclear_f2a_drss (dcs.hh:310-319)
            double a = 1.0;
            if (x < 0.0014)
                a = c.ph_a_low;
            else if (x < 0.04)  // log x = log Q^2 - log(2 M q) in the FP64 kernel
                a = pexp((0.069 * ((Rn ? lQ2 + l_inv_2Mq : nlog<Safe>(x, t)) * INV_LN10) + 0.097) * c.ph_logA, t);
            double poly, dds;
            if constexpr (Safe) {
                poly = 2.0 + x * (-1.85 + x * (2.45 + x * (-2.35 + x)));
                dds = (1. - y + 0.5 * (1. - c.ph_2m2 * inv_Q2) * (y * y + Q2 * inv_EE) * inv_1R) * (inv_Q2 * inv_Q2) -
                      0.25 * inv_EE * inv_Q2;
            } else {
                poly = fma(x, fma(x, fma(x, x - 2.35, 2.45), -1.85), 2.0);
                // 0.5 (1 - 2 m^2 / Q^2): one FMA in the FP64 kernel (the same
                // value: m^2 = 2 m^2 / 2 exactly), as written in the mixed
                // form's recomputation, which then needs one constant less
                const double w = (Rn ? fma(-c.ph_m2, inv_Q2, 0.5) : 0.5 * fma(-c.ph_2m2, inv_Q2, 1.)) *
                                 fma(Q2, inv_EE, yy);
                dds = fma(fma(w, inv_1R, omy), inv_Q2 * inv_Q2, -(q25 * inv_Q2));
            }
            if constexpr (Rn) {  // the per-pair factor A / 2 * 2.603096e-35 / q after the loop
                ds = fma((a * poly) * F2p * (dds * Q2), GL9_W[j], ds);
            } else {
                const double F2A = c.ph_halfA * a * poly * F2p;
                const double d2 = 2.603096E-35 * F2A * dds * inv_q;
                if constexpr (Safe)
                    ds += d2 * Q2 * 1. * GL9_W[j];
                else
                    ds = fma(d2 * Q2, GL9_W[j], ds);
            }
        }
        if constexpr (Rn) ds *= c.ph_halfA * 2.603096E-35 * inv_q;
        return (ds < 0.) ? 0. : dv<!Rn>(0.5 * ds * dpQ2 * 1E+03 * AVOGADRO_NUMBER * (c.mass + K), c.A, c.r_A);
    }

    // The tile runner's interface; FP64 results are final (no recomputation).
    static NOA_DEV_INLINE double pair(double K, double q, const DcsConsts& c, bool&, const fm::Tab& t) {
        return pair(K, q, c, t);
    }
    static NOA_DEV_INLINE double photo(double K, double q, const DcsConsts& c, bool&, const fm::Tab& t) {
        return photo(K, q, c, t);
    }
};

using F64 = F64N<1>;

}  // namespace noa_hip
