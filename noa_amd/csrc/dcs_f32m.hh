// Mixed-precision device physics: FP32 arithmetic and FP32 libm, with FP64
// kept where FP32 would cancel or decide a branch differently from the FP64
// reference:
//   - kinematic differences E - q, K - q and 1 - nu = (E - q) / E;
//   - every exit/threshold test of the reference (done with its FP64
//     arithmetic, so a pair is zero here exactly when it is zero there);
//   - the pair-production integration bound tmin = log(argmin) and the
//     photonuclear Q^2 range log(Q2min), log(Q2max);
//   - the bremsstrahlung screening logarithms near the kinematic limit, where
//     Z Phi_n + Phi_e -> 0 (recomputed in FP64 when small).
// The remaining cancellations are removed algebraically (same mathematics,
// different rounding): 1 - eps = -expm1(t tmin); the non-asymptotic B_e and
// B_mu forms of the pair integrand through h(u) = (log1p(u) - u) / u^2; the
// photonuclear 2Mq - Q^2 = 2M m_pi + m_pi^2 - Q2max expm1(-s), which gives both
// 1 - x and W^2 - M^2 without subtracting nearly equal numbers; and the
// ionisation bracket through Wmax - q and E^2 = P^2 + m^2.
// Formulas: src/noa/pms/physics.hh:114-153, src/noa/pms/dcs.hh:144-443.
#pragma once

#include <hip/hip_runtime.h>

#include "dcs_consts.hh"
#include "dcs_f64.hh"

namespace noa_hip {

static __constant__ float GL8_XF[8] = {0.01985507f, 0.10166676f, 0.2372338f, 0.40828268f,
                                0.59171732f, 0.7627662f, 0.89833324f, 0.98014493f};
static __constant__ float GL8_WF[8] = {0.05061427f, 0.11119052f, 0.15685332f, 0.18134189f,
                                0.18134189f, 0.15685332f, 0.11119052f, 0.05061427f};
// 1 - t_j of the 9-node rule on [-1, 1] (numerics.hh:137-144).
static __constant__ float GL9_OMXF[9] = {1.0f, 1.8360311073266358f, 0.1639688926733642f,
                                  1.9681602395076261f, 0.0318397604923739f, 1.3242534234038089f,
                                  0.6757465765961911f, 1.6133714327005904f, 0.3866285672994096f};
static __constant__ float GL9_WF[9] = {0.3302393550012598f, 0.1806481606948574f, 0.1806481606948574f,
                                0.0812743883615744f, 0.0812743883615744f, 0.3123470770400029f,
                                0.3123470770400029f, 0.2606106964029354f, 0.2606106964029354f};

// a / b through the hardware reciprocal (1 ulp), well inside the 1e-5 budget.
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// log1p(u) for u >= 0: the hardware log2 of y = fl(1 + u) (within 1 ulp of
// log2(y), also as y -> 1) times ln 2, plus the first-order correction
// (u - (y - 1)) / y for the rounding of the sum. Measured on gfx950 over
// u in [1e-7, 2]: within 1.9 FP32 ulp of log1p(u), for five VALU slots and two
// transcendentals instead of the libm log1pf's ~117 instructions.
__device__ __forceinline__ float log1p_f(float u, const fm::Tab&) {
    const float y = 1.0f + u;
    const float c = (u - (y - 1.0f)) * __builtin_amdgcn_rcpf(y);
    return fmaf(__builtin_amdgcn_logf(y), 0.69314718055994531f, c);
}

// expm1(s) in FP32: e^s - 1 from the FP64 table exp has an absolute error
// below 2^-52, under an FP32 ulp of the result for |s| >= 2^-20; smaller |s|
// take s + s^2/2 (truncation s^3/6).
__device__ __forceinline__ float expm1_f(float s, const fm::Tab& t) {
    if (fabsf(s) < 0x1p-20f) return fmaf(0.5f * s, s, s);
    return float(fm::exp_fin(double(s), t) - 1.0);  // s finite: a node offset of the Q^2 range
}

// h(u) = (log1p(u) - u) / u^2 for u > 0, without the cancellation of the
// direct form at small u: below 1/4 a degree-6 polynomial (the Chebyshev
// economization of the alternating series, from gen_fastmath_tables.py;
// < 1e-8 relative before FP32 rounding), the direct form above
// (cancellation factor < 8).
__device__ __forceinline__ float log1p_h(float u, const fm::Tab& t) {
    if (u < 0.25f) {
        float s = fmaf(fm::H_P6, u, fm::H_P5);
        s = fmaf(s, u, fm::H_P4);
        s = fmaf(s, u, fm::H_P3);
        s = fmaf(s, u, fm::H_P2);
        s = fmaf(s, u, fm::H_P1);
        return fmaf(s, u, fm::H_P0);
    }
    return fdiv(log1p_f(u, t) - u, u * u);
}

// NU: unroll factor of the node loops, as for F64N.
template <int NU>
struct F32MN {
    using out_t = float;
    static constexpr bool has_redo = true;

    // Screening sum Z Phi_n + Phi_e of physics.hh:135-148 in FP64.
    static __device__ float brems_screening_f64(double K, double q, const DcsConsts& c, const fm::Tab& t) {
#pragma clang fp contract(off)
        const double me = ELECTRON_MASS;
        const double sqrte = SQRTE_BREMS;
        const double E = K + c.mass;
        const double qe_max = E / (1. + c.b_half_m2 / (me * E));
        const double nu = q / E;
        const double delta = c.b_half_m2 / E * nu / (1. - nu);
        double Phi_n = fm::log(c.b_BZ_n * (c.mass + delta * c.b_Dn_sqrte_m2) / (c.b_D_n * (me + delta * sqrte * c.b_BZ_n)), t);
        if (Phi_n < 0.) Phi_n = 0.;
        double Phi_e = 0.;
        if (q < qe_max) {
            Phi_e = fm::log(c.b_BZe_mass / ((1. + delta * c.b_phie_factor) * (me + delta * sqrte * c.b_BZ_e)), t);
            if (Phi_e < 0.) Phi_e = 0.;
        }
        return float(c.Zd * Phi_n + Phi_e);
    }

    // The range the FP32 arithmetic is laid out for: 1e-20 <= K <= 1e12 GeV,
    // 1e-30 <= q <= 1e12 GeV and q >= 1e-30 E, where the FP32 products and
    // quotients below (E^2, 1 / nu, q^2, ...) stay finite and normal. The
    // tile sends pairs outside it (NaN fails the test) whole to the FP64
    // physics, so the functions below see only pairs inside it.
    static __device__ __forceinline__ bool in_range(double K, double q, const DcsConsts& c) {
        return (K >= 1e-20) && (K <= 1e12) && (q >= 1e-30) && (q <= 1e12) && (q >= 1e-30 * (K + c.mass));
    }

    static __device__ __forceinline__ float brems(double K, double q, const DcsConsts& c, const fm::Tab& t) {
        const float me = float(ELECTRON_MASS);
        const float sqrte = float(SQRTE_BREMS);
        const double E = K + c.mass;
        const float Ef = float(E);
        const float nu = fdiv(float(q), Ef);
        const float omnu = fdiv(float(E - q), Ef);
        const float delta = fdiv(fdiv(float(c.b_half_m2), Ef) * nu, omnu);
        const float BZn = float(c.b_BZ_n), BZe = float(c.b_BZ_e), Dn = float(c.b_D_n);
        float Phi_n = __logf(fdiv(BZn * (float(c.mass) + delta * float(c.b_Dn_sqrte_m2)),
                                  Dn * (me + delta * sqrte * BZn)));
        if (Phi_n < 0.f) Phi_n = 0.f;  // a NaN stays NaN, as in physics.hh:140
        float Phi_e = 0.f;
        // q < E / (1 + m^2 / (2 me E)) without the two divisions
        if (q * (ELECTRON_MASS * E + c.b_half_m2) < ELECTRON_MASS * E * E) {
            Phi_e = __logf(fdiv(float(c.b_BZe_mass), (1.f + delta * float(c.b_phie_factor)) * (me + delta * sqrte * BZe)));
            if (Phi_e < 0.f) Phi_e = 0.f;
        }
        const float Zf = float(c.Zd);
        float S = Zf * Phi_n + Phi_e;
        if (S < 0.05f * (Zf + 1.f)) S = brems_screening_f64(K, q, c, t);
        const float F = 4.f / 3.f * fdiv(omnu, nu) + nu;
        return float(c.f_brems_c) * S * F;
    }

    // The ionisation exits in the reference's FP64 arithmetic (dcs.hh:416-423,
    // no contraction), so that a pair is zero here exactly when it is there.
    static __device__ __forceinline__ bool ion_open(double K, double q, const DcsConsts& c, double& P2, double& E,
                                                    double& Wmax) {
#pragma clang fp contract(off)
        const double me = ELECTRON_MASS;
        P2 = K * (K + c.i_2m);
        E = K + c.mass;
        Wmax = 2. * me * P2 / (c.i_m2 + me * (me + 2. * E));
        if ((Wmax < X_FRACTION * K) || (q > Wmax)) return false;
        return !(q <= c.i_Wmin);
    }

    static __device__ __forceinline__ float ion(double K, double q, const DcsConsts& c, const fm::Tab& t) {
        const double me = ELECTRON_MASS;
        double P2, E, Wmax;
        if (!ion_open(K, q, c, P2, E, Wmax)) return 0.f;
        const float qf = float(q), P2f = float(P2), Ef = float(E), Wf = float(Wmax);
        const float m2f = float(c.i_m2);
        // a0 + (a1 + a2 / q) / q with a1 + a2 / q = ((Wmax - q) / Wmax + m^2 / P^2) / q
        const float inv_q = __builtin_amdgcn_rcpf(qf);
        const float br = fdiv(float(Wmax - q), Wf) + fdiv(m2f, P2f);
        const float cs = float(c.f_ion_c) * Ef * (fdiv(0.5f, P2f) + br * inv_q * inv_q);
        float Delta = 0.f;
        if (K >= c.i_Kthr) {
            const float L1 = log1p_f(2.f * qf / float(me), t);
            Delta = 1.16141E-03f * L1 * (__logf(fdiv(4.f * Ef * float(E - q), m2f)) - L1);
        }
        return cs * (1.f + Delta);
    }

    static __device__ __forceinline__ bool pair_active(double K, double q, const DcsConsts& c) {
        return F64::pair_active_impl<true>(K, q, c);
    }

    // Sets `redo` when the running bound on the FP32 rounding of the two
    // screening logarithms L_e and L_mu (which cancel towards the kinematic
    // edges, e.g. q -> 4 m_e) exceeds PAIR_REDO_TOL of the integral.
    static constexpr float PAIR_REDO_TOL = 2e-6f;

    // (pair and photo see only pairs inside in_range: the tile sends the
    // others whole to the FP64 physics)
    static __device__ float pair(double K, double q, const DcsConsts& c, bool& redo, const fm::Tab& t) {
        // FP64 set-up: the integration bound exactly as dcs.hh:169-176
        double tmin_d;
        {
#pragma clang fp contract(off)
            const double gamma = 1. + K / c.mass;
            const double x0 = 4. * ELECTRON_MASS / q;
            const double x1 = 6. / (gamma * (gamma - q / c.mass));
            const double argmin = (x0 + 2. * (1. - x0) * x1) / (1. + (1. - x1) * sqrt(1. - x0));
            tmin_d = fm::log(argmin, t);
        }
        const double E = K + c.mass;
        const float Ef = float(E);
        const float qf = float(q);
        const float nu = fdiv(qf, Ef);
        const float omnu = fdiv(float(E - q), Ef);
        const float beta = fdiv(0.5f * nu * nu, omnu);
        const float xi_factor = float(c.p_hr2) * beta;
        const float tmin = float(tmin_d);
        const float cL = float(c.p_cL), cLe = float(c.p_cLe), AZ13 = float(c.p_AZ13);
        const float rAZ13 = float(c.p_rAZ13), Z15 = float(c.p_15Z13), inv_r2 = float(1. / c.p_r2);
        const float b2 = 1.f + 2.f * beta;

        float I = 0.f;
        float err = 0.f;  // bound on the absolute FP32 error of I / |tmin|
#pragma unroll NU
        for (int j = 0; j < 8; j++) {
            const float s = GL8_XF[j] * tmin;
            // eps = e^s and rho = 1 - eps from one FP64 table exp
            const double ed = fm::exp_fin(double(s), t);  // s = t_j tmin, finite (tmin from argmin in (0, 1))
            const float eps = float(ed);
            const float rho = (fabsf(s) < 0x1p-20f) ? -fmaf(0.5f * s, s, s) : float(1.0 - ed);
            const float rho2 = rho * rho;
            const float rho21 = eps * (1.f + rho);
            const float xi = xi_factor * rho21;
            const float xi_i = __builtin_amdgcn_rcpf(xi);
            const float opr2 = 1.f + rho2;

            float Be;
            if (xi >= 1E+03f) {
                Be = 0.5f * xi_i * ((3.f - rho2) + 2.f * beta * opr2);
            } else {
                const float l = log1p_f(xi_i, t);
                Be = (2.f + rho2) * (1.f + beta) * l + fdiv(rho21 - beta, 1.f + xi) +
                     (3.f + rho2) * xi_i * log1p_h(xi_i, t);
            }
            const float Ye = fdiv(5.f - rho2 + 4.f * beta * opr2,
                                  2.f * (1.f + 3.f * beta) * __logf(3.f + xi_i) - rho2 - 2.f * beta * (2.f - rho2));
            const float xe = (1.f + xi) * (1.f + Ye);
            const float cLi = fdiv(cL, rho21);
            const float la = __logf(fdiv(AZ13 * sqrtf(xe) * qf, qf + cLi * xe));
            const float lb = log1p_f(cLe * xe, t);
            const float Le = la - 0.5f * lb;
            const float Phi_e = fmaxf(Be * Le, 0.f);
            const float dLe = 3e-7f + 1.2e-7f * (fabsf(la) + fabsf(lb));  // lb: log1p_f, <= 1.9 ulp

            float Bmu;
            if (xi <= 1E-03f) {
                Bmu = 0.5f * xi * (5.f - rho2 + beta * (3.f + rho2));
            } else {
                Bmu = opr2 * (1.f + 1.5f * beta) * log1p_f(xi, t) + fdiv(xi * (rho21 - beta), 1.f + xi) -
                      b2 * rho21 * xi * log1p_h(xi, t);
            }
            const float Ymu = fdiv(4.f + rho2 + 3.f * beta * opr2,
                                   opr2 * (1.5f + 2.f * beta) * __logf(3.f + xi) + 1.f - 1.5f * rho2);
            const float xmu = (1.f + xi) * (1.f + Ymu);
            const float Lmu = __logf(fdiv(rAZ13 * qf, Z15 * (qf + cLi * xmu)));
            const float Phi_mu = fmaxf(Bmu * Lmu, 0.f);
            I = fmaf(-(Phi_e + Phi_mu * inv_r2) * eps * tmin, GL8_WF[j], I);
            const float dLmu = 3e-7f + 1.2e-7f * fabsf(Lmu);
            const float e_node = (Phi_e > 0.f ? fabsf(Be) * dLe : 0.f) + (Phi_mu > 0.f ? fabsf(Bmu) * dLmu * inv_r2 : 0.f);
            err = fmaf(GL8_WF[j] * eps, e_node, err);
        }

        const float gamma = 1.f + float(K / c.mass);
        float zeta = 0.f;
        if (gamma > 35.f) {
            const float Z13 = float(c.p_Z13);
            zeta = 0.073f * __logf(fdiv(gamma, 1.f + float(c.p_g1) * gamma * Z13 * Z13)) - 0.26f;
            if (zeta <= 0.f)
                zeta = 0.f;
            else
                zeta = fdiv(zeta, 0.058f * __logf(fdiv(gamma, 1.f + float(c.p_g2) * gamma * Z13)) - 0.14f);
        }
        const float dcs = float(c.f_pair_c) * (float(c.Zd) + zeta) * fdiv(float(E - q) * I, qf);
        // A NaN integral or result (NaN or infinite inputs that pass the
        // exits) fails the estimate and is recomputed by the FP64 physics,
        // which carries the NaN as dcs.hh:257 does.
        redo = !(err * fabsf(tmin) <= PAIR_REDO_TOL * I) || dcs != dcs;
        return (dcs < 0.f) ? 0.f : dcs;
    }

    static __device__ __forceinline__ bool photo_active(double K, double q, const DcsConsts& c) {
        return F64::photo_active_impl<true>(K, q, c);
    }

    static __device__ float photo(double K, double q, const DcsConsts& c, bool& redo, const fm::Tab& t) {
        const double M = PHOTO_M;
        const double mpi = PHOTO_MPI;
        const double E = K + c.mass;
        double dpQ2_d, Q2max_d, lQ2max_d;
        {
#pragma clang fp contract(off)
            const double y = q / E;
            const double Q2min = c.ph_m2 * y * y / (1. - y);
            Q2max_d = 2.0 * M * (q - mpi) - mpi * mpi;
            lQ2max_d = fm::log(Q2max_d, t);
            dpQ2_d = lQ2max_d - fm::log(Q2min, t);
        }
        const float Ef = float(E);
        const float qf = float(q);
        const float y = fdiv(qf, Ef);
        const float omy = fdiv(float(E - q), Ef);
        const float EE = Ef * Ef;
        const float hd = float(0.5 * dpQ2_d);
        const float Q2max = float(Q2max_d);
        const float c0 = float(2.0 * M * mpi + mpi * mpi);  // 2Mq - Q2max
        const float inv2Mq = float(1.0 / (2.0 * M * q));
        const float L0 = float(c.ph_L0);
        const float logA = float(c.ph_logA);
        const float m2x2 = float(c.ph_2m2);
        const float yy = y * y;

        float ds = 0.f;
#pragma unroll NU
        for (int j = 0; j < 9; j++) {
            const float s = hd * GL9_OMXF[j];  // ln(Q2max / Q2) >= 0
            // Q2 = Q2max e^-s; 2Mq - Q2 = c0 - Q2max expm1(-s) (= W^2 - M^2),
            // with expm1 only where e^-s - 1 would cancel
            const float e = __expf(-s);
            const float em1 = (s < 0.35f) ? expm1_f(-s, t) : e - 1.f;
            const float Q2 = Q2max * e;
            const float D = c0 - Q2max * em1;
            const float x = Q2 * inv2Mq;
            const float omx = D * inv2Mq;
            // ALLM97 F2 (dcs.hh:261-307)
            const float tl = log1p_f(fdiv(log1p_f(fdiv(Q2, float(ALLM_Q02)), t), L0), t);
            const float lnt = __logf(tl);
            const float cP = 0.28067f + (0.28067f - 0.22291f) * (fdiv(1.f, 1.f + __expf(2.1979f * lnt)) - 1.f);
            const float aP = -0.0808f + (-0.0808f + 0.44812f) * (fdiv(1.f, 1.f + __expf(1.1709f * lnt)) - 1.f);
            const float bP = 0.36292f + 1.8917f * __expf(1.8439f * lnt);
            const float cR = 0.80107f + 0.97307f * __expf(3.4942f * lnt);
            const float aR = 0.58400f + 0.37888f * __expf(2.6063f * lnt);
            const float bR = 0.01147f + 3.7582f * __expf(0.49338f * lnt);
            const float xP = fdiv(Q2 + 49.457f, Q2 + 49.457f + D);
            const float xR = fdiv(Q2 + 0.15052f, Q2 + 0.15052f + D);
            const float l1x = __logf(omx);
            const float F2P = cP * __expf(aP * __logf(xP) + bP * l1x);
            const float F2R = cR * __expf(aR * __logf(xR) + bR * l1x);
            const float F2p = fdiv(Q2, Q2 + 0.31985f) * (F2P + F2R);
            // DRSS shadowing (dcs.hh:310-319). Its factor jumps by ~A^5e-4 at
            // x = 0.04, so when FP32 x is too close to call, the branch is
            // decided with the FP64 Bjorken x of this node.
            bool below_hi = x < 0.04f;
            if (fabsf(x - 0.04f) < 1e-7f)
                below_hi = 0.5 * fm::exp(lQ2max_d - 0.5 * dpQ2_d * (1.0 - GL9_X[j]), t) / (M * q) < 0.04;
            float a = 1.f;
            if (x < 0.0014f)
                a = float(c.ph_a_low);
            else if (below_hi)
                a = __expf((0.069f * __log10f(x) + 0.097f) * logA);
            const float F2A = float(c.ph_halfA) * a * (2.f + x * (-1.85f + x * (2.45f + x * (-2.35f + x)))) * F2p;
            // Whitlow R (dcs.hh:322-332)
            const float q2 = fmaxf(Q2, 0.3f);
            const float theta = 1.f + fdiv(12.f * q2, 1.f + q2) * fdiv(0.015625f, 0.015625f + x * x);
            const float R = fdiv(0.635f, __logf(q2 * 25.f)) * theta + fdiv(0.5747f, q2) - fdiv(0.3534f, 0.09f + q2 * q2);
            // dds * Q2 (dcs.hh:348-352)
            const float num = omy + fdiv(0.5f * (1.f - fdiv(m2x2, Q2)) * (yy + fdiv(Q2, EE)), 1.f + R);
            const float ddsQ2 = fdiv(num, Q2) - fdiv(0.25f, EE);
            ds = fmaf(GL9_WF[j], F2A * ddsQ2, ds);
        }
        const float r = float(c.f_photo_c) * float(dpQ2_d) * fdiv(Ef, qf) * ds;
        redo = r != r;  // NaN: recomputed in FP64, which carries it as dcs.hh:404 does
        return (r < 0.f) ? 0.f : r;
    }
};

using F32M = F32MN<1>;

}  // namespace noa_hip
