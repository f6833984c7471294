// CPU oracle for the PMS differential-cross-section (DCS) path.
//
// TEST INFRASTRUCTURE ONLY. This header is the parity checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
// library built from it. The shipped HIP path never calls it.
//
// It restates, in FP64 and in the reference's operation order, the four DCS
// functors of grinisrit/noa:
//   bremsstrahlung   src/noa/pms/physics.hh:114-153 (functor dcs.hh:134-140)
//   pair production  src/noa/pms/dcs.hh:144-258
//   photonuclear     src/noa/pms/dcs.hh:261-405
//   ionisation       src/noa/pms/dcs.hh:408-443
// the fixed-order Gauss-Legendre rules (src/noa/utils/numerics.hh:72-152) and
// the serial / OpenMP element loops (src/noa/utils/common.hh:146-183).
//
// Every expression keeps the reference's evaluation order, so that with the
// same compiler flags (g++ -O3, no FMA contraction on baseline x86-64) the
// results are bit-identical to the compiled reference header. The functions
// are templates over the scalar type T so that an op-counting scalar can
// replay the same arithmetic and report the work per (K, q) pair.
#pragma once

#include <cmath>
#include <cstdint>

namespace noa_oracle {

// physics.hh:54-58,76 and the per-formula literals of dcs.hh.
constexpr double AVOGADRO_NUMBER = 6.02214076E+23;
constexpr double ELECTRON_MASS = 0.510998910E-03;  // GeV/c^2
constexpr double MUON_MASS = 0.10565839;           // GeV/c^2
constexpr double X_FRACTION = 5E-02;

// physics.hh:39-43: A (g/mol), I (GeV), Z.
template <class T>
struct Element {
    T A;
    T I;
    int32_t Z;
};

// numerics.hh:116-121 (quadrature8 literals on [0,1], truncated to 8 decimals;
// these truncated values are part of the specification, not exact GL nodes).
constexpr double GL8_X[8] = {0.01985507, 0.10166676, 0.2372338, 0.40828268,
                             0.59171732, 0.7627662, 0.89833324, 0.98014493};
constexpr double GL8_W[8] = {0.05061427, 0.11119052, 0.15685332, 0.18134189,
                             0.18134189, 0.15685332, 0.11119052, 0.05061427};
// numerics.hh:137-144 (quadrature9 literals on [-1,1], weights sum to 2).
constexpr double GL9_X[9] = {0.0000000000000000, -0.8360311073266358, 0.8360311073266358,
                             -0.9681602395076261, 0.9681602395076261, -0.3242534234038089,
                             0.3242534234038089, -0.6133714327005904, 0.6133714327005904};
constexpr double GL9_W[9] = {0.3302393550012598, 0.1806481606948574, 0.1806481606948574,
                             0.0812743883615744, 0.0812743883615744, 0.3123470770400029,
                             0.3123470770400029, 0.2606106964029354, 0.2606106964029354};
// numerics.hh:97-100 (quadrature6, used by the recoil integrals).
constexpr double GL6_X[6] = {0.03376524, 0.16939531, 0.38069041,
                             0.61930959, 0.83060469, 0.96623476};
constexpr double GL6_W[6] = {0.08566225, 0.18038079, 0.23395697,
                             0.23395697, 0.18038079, 0.08566225};

// numerics.hh:72-89: n_itv = ceil(min_points/order) sub-intervals of width h,
// summed sequentially in node order. With min_points = 1 this is one interval
// of width (ub - lb).
template <class T, class F>
inline T gauss_legendre(const T lb, const T ub, const F& f, uint32_t min_points,
                        uint32_t order, const double* x, const double* w) {
    const uint32_t n_itv = (min_points + order - 1) / order;
    const T h = (ub - lb) / T(double(n_itv));
    const uint32_t n = n_itv * order;
    T acc = T(0.0);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t j = i % order;
        acc += f(lb + h * (T(double(i / order)) + T(x[j]))) * h * T(w[j]);
    }
    return acc;
}

// physics.hh:114-153.
template <class T>
inline T bremsstrahlung(const T K, const T q, const Element<T>& el, const T mass) {
    using std::log;
    using std::pow;
    const int32_t Z = el.Z;
    const T A = el.A;
    const T me = T(ELECTRON_MASS);
    const T sqrte = T(1.648721271);
    const T phie_factor = mass / (me * me * sqrte);
    const T rem = T(5.63588E-13) * me / mass;

    const T BZ_n = (Z == 1) ? T(202.4) : T(182.7) * pow(T(double(Z)), T(-1. / 3.));
    const T BZ_e = (Z == 1) ? T(446.) : T(1429.) * pow(T(double(Z)), T(-2. / 3.));
    const T D_n = T(1.54) * pow(A, T(0.27));
    const T E = K + mass;
    const T dcs_factor = T(7.297182E-07) * rem * rem * T(double(Z));

    const T delta_factor = T(0.5) * mass * mass / E;
    const T qe_max = E / (T(1.) + T(0.5) * mass * mass / (me * E));

    const T nu = q / E;
    const T delta = delta_factor * nu / (T(1.) - nu);
    T Phi_n = log(BZ_n * (mass + delta * (D_n * sqrte - T(2.))) /
                  (D_n * (me + delta * sqrte * BZ_n)));
    if (Phi_n < T(0.)) Phi_n = T(0.);
    T Phi_e;
    if (q < qe_max) {
        Phi_e = log(BZ_e * mass / ((T(1.) + delta * phie_factor) * (me + delta * sqrte * BZ_e)));
        if (Phi_e < T(0.)) Phi_e = T(0.);
    } else {
        Phi_e = T(0.);
    }
    const T dcs = dcs_factor * (T(double(Z)) * Phi_n + Phi_e) *
                  (T(4. / 3.) * (T(1.) / nu - T(1.)) + nu);
    return (dcs < T(0.)) ? T(0.) : dcs * T(1E+03) * T(AVOGADRO_NUMBER) / A;
}

// dcs.hh:179-227, the integrand over t = ln(1 - rho) / tmin on [0, 1].
template <class T>
struct PairIntegrand {
    T tmin, beta, xi_factor, cL, cLe, AZ13, Z13, r, q;
    T operator()(const T t) const {
        using std::log;
        using std::exp;
        using std::sqrt;
        const T eps = exp(t * tmin);
        const T rho = T(1.) - eps;
        const T rho2 = rho * rho;
        const T rho21 = eps * (T(2.) - eps);
        const T xi = xi_factor * rho21;
        const T xi_i = T(1.) / xi;

        T Be;  // electron term, asymptotic form above xi = 1e3
        if (xi >= T(1E+03))
            Be = T(0.5) * xi_i * ((T(3.) - rho2) + T(2.) * beta * (T(1.) + rho2));
        else
            Be = ((T(2.) + rho2) * (T(1.) + beta) + xi * (T(3.) + rho2)) * log(T(1.) + xi_i) +
                 (rho21 - beta) / (T(1.) + xi) - T(3.) - rho2;
        const T Ye = (T(5.) - rho2 + T(4.) * beta * (T(1.) + rho2)) /
                     (T(2.) * (T(1.) + T(3.) * beta) * log(T(3.) + xi_i) - rho2 -
                      T(2.) * beta * (T(2.) - rho2));
        const T xe = (T(1.) + xi) * (T(1.) + Ye);
        const T cLi = cL / rho21;
        const T Le = log(AZ13 * sqrt(xe) * q / (q + cLi * xe)) - T(0.5) * log(T(1.) + cLe * xe);
        T Phi_e = Be * Le;
        if (Phi_e < T(0.)) Phi_e = T(0.);

        T Bmu;  // muon term, asymptotic form below xi = 1e-3
        if (xi <= T(1E-03))
            Bmu = T(0.5) * xi * (T(5.) - rho2 + beta * (T(3.) + rho2));
        else
            Bmu = ((T(1.) + rho2) * (T(1.) + T(1.5) * beta) - xi_i * (T(1.) + T(2.) * beta) * rho21) *
                      log(T(1.) + xi) +
                  xi * (rho21 - beta) / (T(1.) + xi) + (T(1.) + T(2.) * beta) * rho21;
        const T Ymu = (T(4.) + rho2 + T(3.) * beta * (T(1.) + rho2)) /
                      ((T(1.) + rho2) * (T(1.5) + T(2.) * beta) * log(T(3.) + xi) + T(1.) -
                       T(1.5) * rho2);
        const T xmu = (T(1.) + xi) * (T(1.) + Ymu);
        const T Lmu = log(r * AZ13 * q / (T(1.5) * Z13 * (q + cLi * xmu)));
        T Phi_mu = Bmu * Lmu;
        if (Phi_mu < T(0.)) Phi_mu = T(0.);
        return -(Phi_e + Phi_mu / (r * r)) * (T(1.) - rho) * tmin;
    }
};

// dcs.hh:144-258.
template <class T>
inline T pair_production(const T K, const T q, const Element<T>& el, const T mass) {
    using std::log;
    using std::pow;
    using std::sqrt;
    const int32_t Z = el.Z;
    const T A = el.A;
    if (q <= T(4.) * T(ELECTRON_MASS)) return T(0.);
    const T sqrte = T(1.6487212707);
    const T Z13 = pow(T(double(Z)), T(1. / 3.));
    if (q >= K + mass * (T(1.) - T(0.75) * sqrte * Z13)) return T(0.);

    PairIntegrand<T> f;
    const T nu = q / (K + mass);
    f.r = mass / T(ELECTRON_MASS);
    f.beta = T(0.5) * nu * nu / (T(1.) - nu);
    f.xi_factor = T(0.5) * f.r * f.r * f.beta;
    const T A_ = (Z == 1) ? T(202.4) : T(183.);
    f.AZ13 = A_ / Z13;
    f.cL = T(2.) * sqrte * T(ELECTRON_MASS) * f.AZ13;
    f.cLe = T(2.25) * Z13 * Z13 / (f.r * f.r);
    f.Z13 = Z13;
    f.q = q;

    const T gamma = T(1.) + K / mass;
    const T x0 = T(4.) * T(ELECTRON_MASS) / q;
    const T x1 = T(6.) / (gamma * (gamma - q / mass));
    const T argmin = (x0 + T(2.) * (T(1.) - x0) * x1) / (T(1.) + (T(1.) - x1) * sqrt(T(1.) - x0));
    if ((argmin >= T(1.)) || (argmin <= T(0.))) return T(0.);
    f.tmin = log(argmin);

    const T I = gauss_legendre<T>(T(0.), T(1.), f, 1, 8, GL8_X, GL8_W);

    T zeta;  // atomic electrons form factor
    if (gamma <= T(35.)) {
        zeta = T(0.);
    } else {
        T gamma1, gamma2;
        if (Z == 1) {
            gamma1 = T(4.4E-05);
            gamma2 = T(4.8E-05);
        } else {
            gamma1 = T(1.95E-05);
            gamma2 = T(5.30E-05);
        }
        zeta = T(0.073) * log(gamma / (T(1.) + gamma1 * gamma * Z13 * Z13)) - T(0.26);
        if (zeta <= T(0.))
            zeta = T(0.);
        else
            zeta /= T(0.058) * log(gamma / (T(1.) + gamma2 * gamma * Z13)) - T(0.14);
    }

    const T E = K + mass;
    const T dcs = T(1.794664E-34) * T(double(Z)) * (T(double(Z)) + zeta) * (E - q) * I / (q * E);
    return (dcs < T(0.)) ? T(0.) : dcs * T(1E+03) * T(AVOGADRO_NUMBER) * (mass + K) / A;
}

// dcs.hh:261-307, ALLM97 proton structure function F2.
template <class T>
inline T photonuclear_f2_allm(const T x, const T Q2) {
    using std::log;
    using std::exp;
    const T m02 = T(0.31985), mP2 = T(49.457), mR2 = T(0.15052);
    const T Q02 = T(0.52544), Lambda2 = T(0.06527);
    const double cP1 = 0.28067, cP2 = 0.22291, cP3 = 2.1979;
    const double aP1 = -0.0808, aP2 = -0.44812, aP3 = 1.1709;
    const double bP1 = 0.36292, bP2 = 1.8917, bP3 = 1.8439;
    const double cR1 = 0.80107, cR2 = 0.97307, cR3 = 3.4942;
    const double aR1 = 0.58400, aR2 = 0.37888, aR3 = 2.6063;
    const double bR1 = 0.01147, bR2 = 3.7582, bR3 = 0.49338;
    const T M2 = T(0.8803505929);

    const T W2 = M2 + Q2 * (T(1.0) / x - T(1.0));
    const T t = log(log((Q2 + Q02) / Lambda2) / log(Q02 / Lambda2));
    const T xP = (Q2 + mP2) / (Q2 + mP2 + W2 - M2);
    const T xR = (Q2 + mR2) / (Q2 + mR2 + W2 - M2);
    const T lnt = log(t);
    const T cP = T(cP1) + T(cP1 - cP2) * (T(1.0) / (T(1.0) + exp(T(cP3) * lnt)) - T(1.0));
    const T aP = T(aP1) + T(aP1 - aP2) * (T(1.0) / (T(1.0) + exp(T(aP3) * lnt)) - T(1.0));
    const T bP = T(bP1) + T(bP2) * exp(T(bP3) * lnt);
    const T cR = T(cR1) + T(cR2) * exp(T(cR3) * lnt);
    const T aR = T(aR1) + T(aR2) * exp(T(aR3) * lnt);
    const T bR = T(bR1) + T(bR2) * exp(T(bR3) * lnt);

    const T F2P = cP * exp(aP * log(xP) + bP * log(T(1.) - x));
    const T F2R = cR * exp(aR * log(xR) + bR * log(T(1.) - x));
    return Q2 / (Q2 + m02) * (F2P + F2R);
}

// dcs.hh:310-319, DRSS nuclear shadowing applied to F2 of the proton.
template <class T>
inline T photonuclear_f2a_drss(const T x, const T F2p, const T A) {
    using std::log;
    using std::log10;
    using std::exp;
    T a = T(1.0);
    if (x < T(0.0014))
        a = exp(T(-0.1) * log(A));
    else if (x < T(0.04))
        a = exp((T(0.069) * log10(x) + T(0.097)) * log(A));
    return T(0.5) * A * a * (T(2.0) + x * (T(-1.85) + x * (T(2.45) + x * (T(-2.35) + x)))) * F2p;
}

// dcs.hh:322-332, Whitlow R = sigma_L / sigma_T.
template <class T>
inline T photonuclear_r_whitlow(const T x, const T Q2) {
    using std::log;
    T q2 = Q2;
    if (Q2 < T(0.3)) q2 = T(0.3);
    const T theta = T(1.) + T(12.0) * q2 / (T(1.0) + q2) * T(0.015625) / (T(0.015625) + x * x);
    return T(0.635) / log(q2 / T(0.04)) * theta + T(0.5747) / q2 - T(0.3534) / (T(0.09) + q2 * q2);
}

// dcs.hh:335-355, doubly differential cross-section in (nu, Q2).
template <class T>
inline T photonuclear_d2(const T A, const T mass, const T K, const T q, const T Q2) {
    const T cf = T(2.603096E-35);
    const T M = T(0.931494);
    const T E = K + mass;
    const T y = q / E;
    const T x = T(0.5) * Q2 / (M * q);
    const T F2p = photonuclear_f2_allm<T>(x, Q2);
    const T F2A = photonuclear_f2a_drss<T>(x, F2p, A);
    const T R = photonuclear_r_whitlow<T>(x, Q2);
    const T dds = (T(1.) - y + T(0.5) * (T(1.) - T(2.) * mass * mass / Q2) * (y * y + Q2 / (E * E)) /
                                  (T(1.) + R)) /
                      (Q2 * Q2) -
                  T(0.25) / (E * E * Q2);
    return cf * F2A * dds / q;
}

template <class T>
struct PhotonuclearIntegrand {
    T A, mass, K, q, pQ2c, dpQ2;
    T operator()(const T t) const {
        using std::exp;
        const T Q2 = exp(pQ2c + T(0.5) * dpQ2 * t);
        return photonuclear_d2<T>(A, mass, K, q, Q2) * Q2;
    }
};

// dcs.hh:357-405.
template <class T>
inline T photonuclear(const T K, const T q, const Element<T>& el, const T mass) {
    using std::log;
    if ((q < T(1.)) || (q < T(2E-03) * K)) return T(0.);  // dcs.hh:357-359
    const T A = el.A;
    const T M = T(0.931494);
    const T mpi = T(0.134977);
    const T E = K + mass;
    if ((q >= (E - mass)) || (q <= (mpi * (T(1.0) + T(0.5) * mpi / M)))) return T(0.);
    const T y = q / E;
    const T Q2min = mass * mass * y * y / (T(1) - y);
    const T Q2max = T(2.0) * M * (q - mpi) - mpi * mpi;
    if ((Q2max < Q2min) | (Q2min < T(0))) return T(0.);

    const T pQ2min = log(Q2min);
    const T pQ2max = log(Q2max);
    PhotonuclearIntegrand<T> f;
    f.A = A;
    f.mass = mass;
    f.K = K;
    f.q = q;
    f.dpQ2 = pQ2max - pQ2min;
    f.pQ2c = T(0.5) * (pQ2max + pQ2min);
    const T ds = gauss_legendre<T>(T(0.), T(1.), f, 1, 9, GL9_X, GL9_W);
    return (ds < T(0.)) ? T(0.)
                        : T(0.5) * ds * f.dpQ2 * T(1E+03) * T(AVOGADRO_NUMBER) * (mass + K) / A;
}

// dcs.hh:408-443.
template <class T>
inline T ionisation(const T K, const T q, const Element<T>& el, const T mass) {
    using std::log;
    const T A = el.A;
    const int32_t Z = el.Z;
    const T me = T(ELECTRON_MASS);
    const T P2 = K * (K + T(2.) * mass);
    const T E = K + mass;
    const T Wmax = T(2.) * me * P2 / (mass * mass + me * (me + T(2.) * E));
    if ((Wmax < T(X_FRACTION) * K) || (q > Wmax)) return T(0.);
    const T Wmin = T(0.62) * el.I;
    if (q <= Wmin) return T(0.);

    const T a0 = T(0.5) / P2;
    const T a1 = T(-1.) / Wmax;
    const T a2 = E * E / P2;
    const T cs = T(1.535336E-05) * E * T(double(Z)) / A * (a0 + T(1.) / q * (a1 + a2 / q));

    T Delta = T(0.);  // radiative correction
    const T m1 = mass - me;
    if (K >= T(0.5) * m1 * m1 / me) {
        const T L1 = log(T(1.) + T(2.) * q / me);
        Delta = T(1.16141E-03) * L1 * (log(T(4.) * E * (E - q) / (mass * mass)) - L1);
    }
    return cs * (T(1.) + Delta);
}

// ---- recoil integrals (dcs.hh:89-130, 444-497, 955-1001) ----

// dcs.hh:89-105 with del_integrand (dcs * q, :107-109) or cel_integrand
// (dcs * q^2, :111-113): the 6-node rule over t = ln q in [ln(K xlow), ln K]
// with ceil(min_points / 6) subintervals, divided by E = K + mass.
template <class T, class F>
inline T recoil_integral(const F& dcs, bool cel, const T K, const T xlow, const Element<T>& el, const T mass,
                         uint32_t min_points) {
    using std::log;
    using std::exp;
    auto f = [&](const T t) {
        const T q = exp(t);
        const T d = dcs(K, q, el, mass);
        return cel ? d * q * q : d * q;
    };
    return gauss_legendre<T>(log(K * xlow), log(K), f, min_points, 6, GL6_X, GL6_W) / (K + mass);
}

// dcs.hh:469-496 with the close-interaction terms of :446-466.
template <class T>
inline T analytic_ionisation_recoil_integral(const T K, const T xlow, const Element<T>& el, const T mass,
                                             bool cel) {
    using std::log;
    const T me = T(ELECTRON_MASS);
    const T P2 = K * (K + T(2.) * mass);
    const T E = K + mass;
    const T Wmax = T(2.) * me * P2 / (mass * mass + me * (me + T(2.) * E));
    if (Wmax < T(X_FRACTION) * K) return T(0.);
    T Wmin = T(0.62) * el.I;
    const T qlow = K * xlow;
    if (qlow >= Wmin) Wmin = qlow;
    if (Wmax <= Wmin) return T(0.);
    const T a0 = T(0.5) / P2, a1 = T(-1.) / Wmax, a2 = E * E / P2;
    T term;
    if (cel)
        term = T(0.5) * a0 * (Wmax * Wmax - Wmin * Wmin) + a1 * (Wmax - Wmin) + a2 * log(Wmax / Wmin);
    else
        term = a0 * (Wmax - Wmin) + a1 * log(Wmax / Wmin) + a2 * (T(1.) / Wmin - T(1.) / Wmax);
    return T(1.535336E-05) * T(double(el.Z)) / el.A * term;
}

// The recoil integral of one process as the reference dispatches it: the
// ionisation specialisations (dcs.hh:955-1001) take the analytic form for
// K <= (m - m_e)^2 / (2 m_e) and the numerical one above.
template <class T>
inline T process_recoil_integral(int process, bool cel, const T K, const T xlow, const Element<T>& el,
                                 const T mass, uint32_t min_points) {
    switch (process) {
        case 0: return recoil_integral<T>(bremsstrahlung<T>, cel, K, xlow, el, mass, min_points);
        case 1: return recoil_integral<T>(pair_production<T>, cel, K, xlow, el, mass, min_points);
        case 2: return recoil_integral<T>(photonuclear<T>, cel, K, xlow, el, mass, min_points);
        default: {
            const T m1 = mass - T(ELECTRON_MASS);
            if (K <= T(0.5) * m1 * m1 / T(ELECTRON_MASS))
                return analytic_ionisation_recoil_integral<T>(K, xlow, el, mass, cel);
            return recoil_integral<T>(ionisation<T>, cel, K, xlow, el, mass, min_points);
        }
    }
}

}  // namespace noa_oracle
