// Per-launch constants of one (element, projectile mass) pair.
//
// The reference recomputes these inside every scalar call (pow(Z, -1/3),
// pow(A, 0.27), exp(-0.1 log A), ...; src/noa/pms/physics.hh:119-133,
// src/noa/pms/dcs.hh:148-166, 312-313, 415-436). Here they are computed once
// on the host, with the host libm and the same operation order, so every
// hoisted value is bit-identical to the one the reference's CPU path uses,
// and the kernels read them from the kernarg segment (scalar loads).
#pragma once

#include <cmath>
#include <cstdint>

namespace noa_hip {

constexpr double AVOGADRO_NUMBER = 6.02214076E+23;  // physics.hh:54
constexpr double ELECTRON_MASS = 0.510998910E-03;   // physics.hh:57
constexpr double X_FRACTION = 5E-02;                // physics.hh:76
constexpr double SQRTE_BREMS = 1.648721271;         // physics.hh:122
constexpr double SQRTE_PAIR = 1.6487212707;         // dcs.hh:153
constexpr double PHOTO_M = 0.931494;                // dcs.hh:339, 370
constexpr double PHOTO_MPI = 0.134977;              // dcs.hh:371
constexpr double ALLM_Q02 = 0.52544;                // dcs.hh:265
constexpr double ALLM_LAMBDA2 = 0.06527;            // dcs.hh:266

struct DcsConsts {
    // element and projectile
    double A, I, Zd, mass;
    int32_t Z, z_is_1;
    // bremsstrahlung
    double b_BZ_n, b_BZ_e, b_D_n, b_phie_factor, b_dcs_factor;
    double b_half_m2;        // 0.5 * mass * mass
    double b_Dn_sqrte_m2;    // D_n * sqrte - 2
    double b_BZe_mass;       // BZ_e * mass
    // pair production
    double p_Z13, p_cut;     // Z^(1/3), mass * (1 - 0.75 sqrte Z^(1/3))
    double p_r, p_hr2;       // mass / me, 0.5 * r * r
    double p_AZ13, p_cL, p_cLe, p_rAZ13, p_15Z13, p_r2, p_inv_r2, p_AZ13_sq;
    double p_g1, p_g2, p_dcs_c;  // zeta gammas, 1.794664E-34 * Z
    // photonuclear
    double ph_logA, ph_a_low, ph_halfA, ph_2m2, ph_m2, ph_thr_pi, ph_L0, ph_inv_L0;
    // ionisation
    double i_2m, i_m2, i_Wmin, i_Kthr;
    // folded output prefactors of the mixed-precision form (FP64 here,
    // rounded once to FP32 in the kernel; each keeps FP32 clear of underflow)
    double f_brems_c;  // dcs_factor * 1e3 * N_A / A
    double f_pair_c;   // 1.794664e-34 * Z * 1e3 * N_A / A
    double f_photo_c;  // 2.603096e-35 * 0.5 * 1e3 * N_A / A
    double f_ion_c;    // 1.535336e-5 * Z / A
    // correctly rounded reciprocals for the quotients of the FP64 closed
    // forms (the kernels divide through them with a Markstein step)
    double r_A, r_mass, i_r_m2;  // 1 / A, 1 / mass, 1 / mass^2
    // 1 when mass and A lie where every closed-form quotient of an ordinary
    // (K, q) stays in the normal range (1e-6 <= mass <= 1e6 GeV, 1 <= A <=
    // 1e4 g/mol); otherwise the kernels keep the reference's divisions
    int32_t fast_ok, pad_;
};

// The same arithmetic, in the same order, as the reference's inline
// computations; evaluated with the host libm.
inline DcsConsts make_consts(double A, double I, int32_t Z, double mass) {
    DcsConsts c{};
    const double me = ELECTRON_MASS;
    c.A = A;
    c.I = I;
    c.Z = Z;
    c.Zd = double(Z);
    c.z_is_1 = (Z == 1);
    c.mass = mass;
    // physics.hh:121-132
    {
        const double sqrte = SQRTE_BREMS;
        c.b_phie_factor = mass / (me * me * sqrte);
        const double rem = 5.63588E-13 * me / mass;
        c.b_BZ_n = (Z == 1) ? 202.4 : 182.7 * std::pow(double(Z), -1. / 3.);
        c.b_BZ_e = (Z == 1) ? 446. : 1429. * std::pow(double(Z), -2. / 3.);
        c.b_D_n = 1.54 * std::pow(A, 0.27);
        c.b_dcs_factor = 7.297182E-07 * rem * rem * double(Z);
        c.b_half_m2 = 0.5 * mass * mass;
        c.b_Dn_sqrte_m2 = c.b_D_n * sqrte - 2.;
        c.b_BZe_mass = c.b_BZ_e * mass;
    }
    // dcs.hh:153-166, 236-255
    {
        const double sqrte = SQRTE_PAIR;
        c.p_Z13 = std::pow(double(Z), 1. / 3.);
        c.p_cut = mass * (1. - 0.75 * sqrte * c.p_Z13);
        c.p_r = mass / me;
        c.p_hr2 = 0.5 * c.p_r * c.p_r;
        const double A_ = (Z == 1) ? 202.4 : 183.;
        c.p_AZ13 = A_ / c.p_Z13;
        c.p_cL = 2. * sqrte * me * c.p_AZ13;
        c.p_cLe = 2.25 * c.p_Z13 * c.p_Z13 / (c.p_r * c.p_r);
        c.p_rAZ13 = c.p_r * c.p_AZ13;
        c.p_15Z13 = 1.5 * c.p_Z13;
        c.p_r2 = c.p_r * c.p_r;
        c.p_inv_r2 = 1. / c.p_r2;
        c.p_AZ13_sq = c.p_AZ13 * c.p_AZ13;
        c.p_g1 = (Z == 1) ? 4.4E-05 : 1.95E-05;
        c.p_g2 = (Z == 1) ? 4.8E-05 : 5.30E-05;
        c.p_dcs_c = 1.794664E-34 * double(Z);
    }
    // dcs.hh:290, 313, 317, 349, 374
    {
        c.ph_logA = std::log(A);
        c.ph_a_low = std::exp(-0.1 * c.ph_logA);
        c.ph_halfA = 0.5 * A;
        c.ph_2m2 = 2. * mass * mass;
        c.ph_m2 = mass * mass;
        c.ph_thr_pi = PHOTO_MPI * (1.0 + 0.5 * PHOTO_MPI / PHOTO_M);
        volatile double q0 = ALLM_Q02, l2 = ALLM_LAMBDA2;  // runtime libm, not folded
        c.ph_L0 = std::log(q0 / l2);
        c.ph_inv_L0 = 1. / c.ph_L0;
    }
    // dcs.hh:415-436
    {
        c.i_2m = 2. * mass;
        c.i_m2 = mass * mass;
        c.i_Wmin = 0.62 * I;
        const double m1 = mass - me;
        c.i_Kthr = 0.5 * m1 * m1 / me;
    }
    c.f_brems_c = c.b_dcs_factor * 1E+03 * AVOGADRO_NUMBER / A;
    c.f_pair_c = c.p_dcs_c * 1E+03 * AVOGADRO_NUMBER / A;
    c.f_photo_c = 2.603096E-35 * 0.5 * 1E+03 * AVOGADRO_NUMBER / A;
    c.f_ion_c = 1.535336E-05 * double(Z) / A;
    c.r_A = 1. / A;
    c.r_mass = 1. / mass;
    c.i_r_m2 = 1. / c.i_m2;
    c.fast_ok = (mass >= 1e-6 && mass <= 1e6 && A >= 1. && A <= 1e4) ? 1 : 0;
    return c;
}

}  // namespace noa_hip
