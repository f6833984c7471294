// CPU oracle for the Coulomb-scattering helpers of the PMS DCS path
// (TEST INFRASTRUCTURE ONLY: loaded by tests/ through liboracle.so).
//
// FP64 restatement, in the reference's operation order, of
//   frame / spin / Wentzel path    src/noa/pms/dcs.hh:499-537
//   screening parameters           src/noa/pms/dcs.hh:540-602
//   coulomb_data loop              src/noa/pms/dcs.hh:604-622
//   transport coefficients         src/noa/pms/dcs.hh:624-694
//   restricted cross section       src/noa/pms/dcs.hh:697-735
//   cutoff objective, hard scatter src/noa/pms/dcs.hh:737-872
//   Ridders' root                  src/noa/utils/numerics.hh:154-214
//   transverse transport (soft)    src/noa/pms/dcs.hh:874-950
// with the constants of src/noa/pms/physics.hh:55,78-87. Layouts follow the
// reference's tensors: fCM [n][2], screening [n][9], G [n][2]; for a mixture
// of nel elements the per-element arrays are [nel][nkin][...].
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "dcs_oracle.hh"

namespace noa_oracle {

constexpr double ATOMIC_MASS_ENERGY = 0.931494;  // physics.hh:55
constexpr double KIN_CUTOFF = 1E-9;              // physics.hh:78
constexpr double EHS_PATH_MAX = 1E+9;            // physics.hh:79
constexpr double EHS_OVER_MSC = 1E-4;            // physics.hh:80
constexpr double MAX_SOFT_ANGLE = 1E+00;         // physics.hh:81, degrees
constexpr int NSF = 9;                           // physics.hh:87

// physics.hh:83-84, evaluated at run time as the reference does.
inline double max_mu0() { return 0.5 * (1. - std::cos(MAX_SOFT_ANGLE * M_PI / 180.)); }

// dcs.hh:499-521: Lorentz factors to the centre-of-mass frame; returns the
// kinetic energy in that frame, floored at KIN_CUTOFF.
inline double coulomb_frame_parameters(double* fCM, double K, const Element<double>& el, double mass) {
    const double Ma = el.A * ATOMIC_MASS_ENERGY;
    double M2 = mass + Ma;
    M2 *= M2;
    const double s_inv = 1. / std::sqrt(M2 + 2. * Ma * K);
    fCM[0] = (K + mass + Ma) * s_inv;
    double k0 = (K * Ma + mass * (mass + Ma)) * s_inv - mass;
    if (k0 < KIN_CUTOFF) k0 = KIN_CUTOFF;
    const double etot = K + mass + Ma;
    const double beta2 = K * (K + 2. * mass) / (etot * etot);
    double rM2 = mass / Ma;
    rM2 *= rM2;
    fCM[1] = std::sqrt(rM2 * (1. - beta2) + beta2);
    return k0;
}

// dcs.hh:524-527.
inline double coulomb_spin_factor(double K, double mass) {
    const double e = K + mass;
    return K * (e + mass) / (e * e);
}

// dcs.hh:530-537.
inline double coulomb_wentzel_path(double screening, double K, const Element<double>& el, double mass) {
    const double d = K * (K + 2. * mass) / (el.Z * (K + mass));
    return el.A * 2.54910918E+08 * screening * (1. + screening) * d * d;
}

// dcs.hh:540-602: atomic (Moliere with Kuraev's Coulomb correction) and two
// nuclear screening parameters, the pole-reduction coefficients [3..8], and
// the inverse Wentzel path as the return value.
inline double coulomb_screening_parameters(double* s, double K, const Element<double>& el, double mass) {
    const double A13 = std::pow(el.A, 1. / 3);
    const double R1 = 1.02934 * A13 + 0.435;
    const double R2 = 2.;
    const double p2 = K * (K + 2. * mass);
    const double d = 5.8406E-02 / p2;
    s[1] = d / (R1 * R1);
    s[2] = d / (R2 * R2);

    const int32_t Z = el.Z;
    const double etot = K + mass;
    const double ZE = Z * etot;
    const double zeta2 = 5.3251346E-05 * (ZE * ZE) / p2;
    double cK;
    if (zeta2 > 1.) {
        const int32_t terms = 10 + Z;
        double f = 0.;
        for (int32_t i = 1; i <= terms; i++) f += zeta2 / (i * (i * i + zeta2));
        cK = std::exp(f);
    } else {
        cK = std::exp(1. - 1. / (1. + zeta2) + zeta2 * (0.2021 + zeta2 * (0.0083 * zeta2 - 0.0369)));
    }
    const double cM = 1. + 3.34 * zeta2;
    double r = K / etot;
    r *= r;
    const double c = r * cK + (1. - r) * cM;
    s[0] = 5.179587126E-12 * std::pow(double(Z), 2. / 3.) * c / p2;

    const double d01 = 1. / (s[0] - s[1]);
    const double d02 = 1. / (s[0] - s[2]);
    const double d12 = 1. / (s[1] - s[2]);
    s[6] = d01 * d01 * d02 * d02;
    s[7] = d01 * d01 * d12 * d12;
    s[8] = d12 * d12 * d02 * d02;
    s[3] = 2. * s[6] * (d01 + d02);
    s[4] = 2. * s[7] * (d12 - d01);
    s[5] = -2. * s[8] * (d12 + d02);
    return 1. / coulomb_wentzel_path(s[0], K, el, mass);
}

// dcs.hh:604-622, one kinetic energy.
inline void coulomb_data(double* fCM, double* screen, double* fspin, double* invlambda, double K,
                         const Element<double>& el, double mass) {
    const double k0 = coulomb_frame_parameters(fCM, K, el, mass);
    *fspin = coulomb_spin_factor(k0, mass);
    *invlambda = coulomb_screening_parameters(screen, k0, el, mass);
}

// dcs.hh:624-674: first and second transport coefficients for the angular
// cutoff mu.
inline void coulomb_transport_coefficients(double* G, const double* s, double fspin, double mu) {
    const double nuc = (s[1] < s[2]) ? s[1] : s[2];
    if (mu < 1E-08 * nuc) {
        const double L = std::log(1. + mu / s[0]);
        const double r = mu / (mu + s[0]);
        const double k = s[0] * (1. + s[0]);
        G[0] = k * (r / s[0] - fspin * (L - r));
        const double I2 = mu - s[0] * (r - 2. * L);
        G[1] = 2. * k * (L - r - fspin * I2);
        return;
    }
    double I0[3], I1[3], I2[3], J0[3], J1[3], J2[3];
    const double half_mu2 = 0.5 * mu * mu;
    for (int i = 0; i < 3; i++) {
        double r = mu / (mu + s[i]);
        double L = std::log(1. + mu / s[i]);
        double mu1 = mu;
        I0[i] = r / s[i];
        J0[i] = L;
        I1[i] = L - r;
        r *= s[i];
        L *= s[i];
        J1[i] = mu1 - L;
        I2[i] = mu1 - 2. * L + r;
        L *= s[i];
        mu1 *= s[i];
        J2[i] = half_mu2 + L - mu1;
    }
    const double k = s[0] * (1. + s[0]) * s[1] * s[1] * s[2] * s[2];
    double g0 = 0., g1 = 0.;
    for (int i = 0; i < 3; i++) {
        g0 += s[3 + i] * (J0[i] - fspin * J1[i]) + s[6 + i] * (I0[i] - fspin * I1[i]);
        g1 += s[3 + i] * (J1[i] - fspin * J2[i]) + s[6 + i] * (I1[i] - fspin * I2[i]);
    }
    G[0] = g0 * k;
    G[1] = g1 * (2. * k);
}

// dcs.hh:697-735: cross section restricted to deflections mu' > mu.
inline double coulomb_restricted_cs(double mu, double fspin, const double* s) {
    if (mu >= 1.) return 0.;
    const double nuc = (s[1] < s[2]) ? s[1] : s[2];
    if (mu < 1E-08 * nuc) {
        const double L = std::log((s[0] + 1.) / (s[0] + mu));
        const double r = (1. - mu) / ((s[0] + mu) * (s[0] + 1.));
        const double k = s[0] * (1. + s[0]);
        return k * (r - fspin * (L - s[0] * r));
    }
    double I0[3], I1[3], J0[3], J1[3];
    for (int i = 0; i < 3; i++) {
        const double L = std::log((s[i] + 1.) / (s[i] + mu));
        const double r = (1. - mu) / ((s[i] + mu) * (s[i] + 1.));
        I0[i] = r;
        J0[i] = L;
        I1[i] = L - s[i] * r;
        J1[i] = mu - s[i] * L;
    }
    const double k = s[0] * (1. + s[0]) * s[1] * s[1] * s[2] * s[2];
    double cs = 0.;
    for (int i = 0; i < 3; i++) cs += s[3 + i] * (J0[i] - fspin * J1[i]) + s[6 + i] * (I0[i] - fspin * I1[i]);
    return k * cs;
}

// Per-kinetic-energy view of a mixture: element e of K index i sits at
// offset e * nkin + i of the per-element arrays.
struct MixtureView {
    const double* G;          // [nel][nkin][2]
    const double* fCM;        // [nel][nkin][2]
    const double* screen;     // [nel][nkin][9]
    const double* invlambda;  // [nel][nkin]
    const double* fspin;      // [nel][nkin]
    int64_t nel, nkin, i;
    int64_t off(int64_t e) const { return e * nkin + i; }
};

// dcs.hh:737-751.
inline double cutoff_objective(double cs_h, double mu, const MixtureView& m) {
    double cs = 0.;
    for (int64_t e = 0; e < m.nel; e++) {
        const int64_t o = m.off(e);
        cs += m.invlambda[o] * coulomb_restricted_cs(mu, m.fspin[o], m.screen + NSF * o);
    }
    return cs - cs_h;
}

// numerics.hh:154-214 (Ridders' method with given end values); returns
// false where the reference returns std::nullopt.
template <class F>
inline bool ridders_root(double xa, double xb, const F& f, double fa, double fb, double xtol, double rtol,
                         uint32_t max_iter, double* root) {
    if (fa * fb > 0) return false;
    if (fa == 0) { *root = xa; return true; }
    if (fb == 0) { *root = xb; return true; }
    const double tol = xtol + rtol * std::min(std::abs(xa), std::abs(xb));
    for (uint32_t it = 0; it < max_iter; it++) {
        double dm = 0.5 * (xb - xa);
        const double xm = xa + dm;
        const double fm = f(xm);
        double sgn = (fb > fa) ? 1. : -1.;
        double dn = sgn * dm * fm / std::sqrt(fm * fm - fa * fb);
        sgn = (dn > 0.) ? 1. : -1.;
        dn = std::abs(dn);
        dm = std::abs(dm) - 0.5 * tol;
        if (dn < dm) dm = dn;
        const double xn = xm - sgn * dm;
        const double fn = f(xn);
        if (fn * fm < 0.0) {
            xa = xn; fa = fn; xb = xm; fb = fm;
        } else if (fn * fa < 0.0) {
            xb = xn; fb = fn;
        } else {
            xa = xn; fa = fn;
        }
        if (fn == 0.0 || std::abs(xb - xa) < tol) { *root = xn; return true; }
    }
    return false;
}

// dcs.hh:753-838: hard-scattering cutoff angle mu0 and mean free path lb_h.
inline void coulomb_hard_scattering(double* mu0_out, double* lb_h_out, const MixtureView& m) {
    double invlb_m = 0., invlb1_m = 0., s_m_l = 0., s_m_h = 0.;
    for (int64_t e = 0; e < m.nel; e++) {
        const int64_t o = m.off(e);
        const double invlb = m.invlambda[o];
        const double scr = m.screen[NSF * o];
        invlb_m += invlb * m.G[2 * o];
        s_m_h += scr * invlb;
        s_m_l += invlb / scr;
        const double d = 1. / (m.fCM[2 * o] * (1. + m.fCM[1 + 2 * o]));
        invlb1_m += invlb * m.G[1 + 2 * o] * d * d;
    }
    const double lb_m = 1. / invlb_m;
    double lb_h = std::min(EHS_OVER_MSC / invlb1_m, EHS_PATH_MAX);
    double mu0;
    if (lb_m < lb_h) {
        const double s_m = (lb_h > 2. * lb_m) ? s_m_h * lb_m : 1. / (s_m_l * lb_m);
        mu0 = s_m * (lb_h - lb_m) / (s_m * lb_h + lb_m);
        const double cs_h = 1. / lb_h;
        auto obj = [&](double mu) { return cutoff_objective(cs_h, mu, m); };
        double mu_max = std::min(4. * mu0, 1.);
        double mu_min = 0.25 * mu0;
        double fmax = obj(mu_max), fmin;
        if (fmax > 0.) {
            // the reference leaves mu0 and lb_h at their estimates here
            mu_min = mu_max;
            fmin = fmax;
            mu_max = 1.;
            fmax = -cs_h;
        } else {
            fmin = obj(mu_min);
            if (fmin < 0.) {
                mu_max = mu_min;
                fmax = fmin;
                mu_min = 0.;
                fmin = obj(mu_min);
            }
            const double mmax = max_mu0();
            if (mu_min < mmax) {
                mu_max = std::min(mu_max, mmax);
                double root;
                if (ridders_root(mu_min, mu_max, obj, fmin, fmax, 1E-6 * mu0, 1E-6, 100, &root)) mu0 = root;
            }
            mu0 = std::min(mu0, mmax);
            lb_h = obj(mu0) + cs_h;
            lb_h = (lb_h <= 1. / EHS_PATH_MAX) ? EHS_PATH_MAX : 1. / lb_h;
        }
    } else {
        lb_h = lb_m;
        mu0 = 0;
    }
    *mu0_out = mu0;
    *lb_h_out = lb_h;
}

// dcs.hh:874-897: soft close collisions below X_FRACTION.
inline double transverse_transport_ionisation(double K, const Element<double>& el, double mass) {
    const double p2 = K * (K + 2. * mass);
    const double E = K + mass;
    const double Wmax = 2. * ELECTRON_MASS * p2 / (mass * mass + ELECTRON_MASS * (ELECTRON_MASS + 2. * E));
    const double W0 = 2. * p2 / ELECTRON_MASS;
    const double mu_max = Wmax / W0;
    double mu3 = K * X_FRACTION / W0;
    if (mu3 > mu_max) mu3 = mu_max;
    const double mu2 = 0.62 * el.I / W0;
    if (mu2 >= mu3) return 0.;
    const double a0 = 0.5 * W0 / p2;
    const double a1 = -1. / Wmax;
    const double a2 = E * E / (W0 * p2);
    const double cs0 = 1.535336E-05 / el.A;
    return 2. * cs0 * el.Z * (0.5 * a0 * (mu3 * mu3 - mu2 * mu2) + a1 * (mu3 - mu2) + a2 * std::log(mu3 / mu2));
}

// dcs.hh:901-938, the mu-integrated ratio at one node of nu.
inline double photonuclear_transport_ratio(double K, double q, double mass) {
    const double E = K + mass;
    const double m02 = 0.4;
    const double q2 = q * q;
    const double tmax = 1.876544 * q;
    const double tmin = q2 * mass * mass / (E * (E - q));
    const double b1 = 1. / (1. - q2 / m02);
    const double c1 = 1. / (1. - m02 / q2);
    double L1 = b1 * std::log((q2 + tmax) / (q2 + tmin));
    double L2 = c1 * std::log((m02 + tmax) / (m02 + tmin));
    const double I0 = std::log(tmax / tmin) - L1 - L2;
    L1 *= q2;
    L2 *= m02;
    const double I1 = L1 + L2;
    L1 *= q2;
    L2 *= m02;
    const double I2 = (tmax - tmin) * (b1 * q2 + c1 * m02) - L1 - L2;
    return (I1 * tmax - I2) / ((I0 * tmax - I1) * K * (K + 2. * mass));
}

// dcs.hh:901-938: 2 x the 6-node rule with 100 minimum points over
// t = ln(nu / X_FRACTION) in [ln 1e-6, 0] of photonuclear * ratio * nu.
inline double transverse_transport_photonuclear(double K, const Element<double>& el, double mass) {
    auto f = [&](const double t) {
        const double nu = X_FRACTION * std::exp(t);
        const double q = nu * K;
        const double ratio = photonuclear_transport_ratio(K, q, mass);
        return photonuclear<double>(K, q, el, mass) * ratio * nu;
    };
    return 2. * gauss_legendre<double>(std::log(1E-06), 0., f, 100, 6, GL6_X, GL6_W);
}

// dcs.hh:940-950.
inline double soft_scattering(double K, const Element<double>& el, double mass) {
    return transverse_transport_ionisation(K, el, mass) + transverse_transport_photonuclear(K, el, mass);
}

}  // namespace noa_oracle
