// CPU oracle (test infrastructure only) of the PUMAS v1.2.1 DCS tabulation
// vendored by the reference (src/noa/3rdparty/_pumas/pumas.c): the default
// DCS models (SSR bremsstrahlung, SSR pair production, DRSS photonuclear,
// pumas.c:190-198), their m^2/kg wrappers, the x grid and the per-row table
// of compute_dcs_table / dcs_tabulate_row. A restatement in the reference's
// operation order with glibc's libm, so that its values equal the compiled
// reference's bit for bit (pinned by tests/test_oracle_pumas.py against
// tests/golden/pumas_golden.npz, written by oracle/_ref/libpumas_ref.so).
// The envelope maximum (dcs_tabulate_envelope_row, pumas.c:11509-11679) is
// restated with its bisection searches (math_find_minimum with algo 0,
// pumas.c:12172-12232, 12416-12459) and its table interpolation
// (dcs_evaluate, pumas.c:11307-11398).
// Only the tests, the smoke check and the benchmark's CPU leg use it.
#pragma once

#include <cmath>
#include <cstdint>

namespace noa_oracle {
namespace pumas {

// pumas.c:128-186
constexpr double ALPHA_EM = 7.2973525693E-03;
constexpr double HBAR_C = 1.973269804E-16;
constexpr double ELECTRON_MASS = 0.510998910E-03;
constexpr double PROTON_MASS = 0.938272;
constexpr double NEUTRON_MASS = 0.939565;
constexpr double PION_MASS = 0.13957018;
constexpr double AVOGADRO_NUMBER = 6.02214076E+23;
constexpr double PI = 3.14159265358979323846;
// pumas.c:109-119
constexpr int DCS_MODEL_N_MIN = 10;
constexpr int DCS_MODEL_PER_DECADE = 15;
constexpr int DCS_MODEL_N_REVERSE = 20;
constexpr double DCS_MODEL_X_REVERSE = 0.6;
constexpr double DCS_MODEL_DXMIN = 1E-03;
constexpr double DCS_MODEL_MAX_FRACTION = 0.95;

// pumas.c:12527-12614: Gauss-Legendre nodes and weights on [0, 1] for 9 and
// 12 points (the only orders the tabulated models use).
constexpr double GQ9_X[9] = {0.0159198802461870, 0.0819844463366821, 0.1933142836497048, 0.3378732882980955,
                             0.5000000000000000, 0.6621267117019045, 0.8066857163502952, 0.9180155536633179,
                             0.9840801197538130};
constexpr double GQ9_W[9] = {0.0406371941807872, 0.0903240803474287, 0.1303053482014677, 0.1561735385200015,
                             0.1651196775006299, 0.1561735385200015, 0.1303053482014677, 0.0903240803474287,
                             0.0406371941807872};
constexpr double GQ12_X[12] = {0.0092196828766404, 0.0479413718147625, 0.1150486629028477, 0.2063410228566913,
                               0.3160842505009099, 0.4373832957442655, 0.5626167042557344, 0.6839157494990901,
                               0.7936589771433087, 0.8849513370971523, 0.9520586281852375, 0.9907803171233596};
constexpr double GQ12_W[12] = {0.0235876681932559, 0.0534696629976592, 0.0800391642716731, 0.1015837133615330,
                               0.1167462682691774, 0.1245735229067014, 0.1245735229067014, 0.1167462682691774,
                               0.1015837133615330, 0.0800391642716731, 0.0534696629976592, 0.0235876681932559};

// pumas.c:13773-13802
inline double radiation_logarithm(double Z) {
    const int i = (int)Z;
    if ((i >= 1) && (i <= 92)) {
        static const double Lz[92] = {
            202.4, 151.9, 159.9, 172.3, 177.9, 178.3, 176.6, 173.4, 170.0, 165.8, 165.8, 167.1, 169.1, 170.8,
            172.2, 173.4, 174.3, 174.8, 175.1, 175.6, 176.2, 176.8, 0.0,   0.0,   0.0,   175.8, 0.0,   0.0,
            173.1, 0.0,   0.0,   173.0, 0.0,   0.0,   173.5, 0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   175.9,
            0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   177.4, 0.0,   0.0,   178.6, 0.0,   0.0,   0.0,
            0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,
            0.0,   0.0,   0.0,   177.6, 0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   178.0, 0.0,   0.0,
            0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   179.8};
        const double l = Lz[i - 1];
        if (l > 0) return l;
    }
    return 182.7;
}

// pumas.c:12673-12738 (CERNLIB RDILOG, Chebyshev series).
inline double dilog(double x) {
    static const double C[20] = {
        0.42996693560813697,  0.40975987533077105,  -0.01858843665014592, 0.00145751084062268,
        -0.00014304184442340, 0.00001588415541880,  -0.00000190784959387, 0.00000024195180854,
        -0.00000003193341274, 0.00000000434545063,  -0.00000000060578480, 0.00000000008612098,
        -0.00000000001244332, 0.00000000000182256,  -0.00000000000027007, 0.00000000000004042,
        -0.00000000000000610, 0.00000000000000093,  -0.00000000000000014, 0.00000000000000002};
    const double PI3 = PI * PI / 3.;
    const double PI6 = PI * PI / 6.;
    const double PI12 = PI * PI / 12.;
    if (x == 10) return PI6;
    if ((x == -1.) || (1. - x == 0.)) return -PI12;
    const double T = -x;
    double A, S, Y;
    if (T <= -2) {
        Y = -1. / (1. + T);
        S = 1.;
        A = -PI3 + 0.5 * (std::pow(std::log(-T), 2.) - std::pow(std::log(1. + 1. / T), 2.));
    } else if (T < -1.) {
        Y = -1. - T;
        S = -1.;
        A = std::log(-T);
        A = -PI6 + A * (A + std::log(1. + 1. / T));
    } else if (T <= -0.5) {
        Y = -(1. + T) / T;
        S = 1.;
        A = std::log(-T);
        A = -PI6 + A * (-0.5 * A + std::log(1. + T));
    } else if (T < 0.) {
        Y = -T / (1. + T);
        S = -1.;
        A = 0.5 * std::pow(std::log(1 + T), 2.);
    } else if (T <= 1.) {
        Y = T;
        S = 1.;
        A = 0.;
    } else {
        Y = 1. / T;
        S = -1.;
        A = PI6 + 0.5 * std::pow(std::log(T), 2);
    }
    const double H = Y + Y - 1;
    const double ALFA = H + H;
    double B0 = 0., B1 = 0., B2 = 0.;
    for (int i = 19; i >= 0; i--) {
        B0 = C[i] + ALFA * B1 - B2;
        B2 = B1;
        B1 = B0;
    }
    return -(S * (B0 - H * B2) + A);
}

// pumas.c:13979-14068, atomic DCS in m^2 / GeV.
inline double brems_SSR(double Z, double A, double m, double K, double q) {
    const double SQRTE = 1.648721270700128, ME = 0.5109989461, RE = 2.8179403227E-13, MMU = 105.6583745;
    if ((Z <= 0) || (A <= 0) || (m <= 0) || (K <= 0) || (q <= 0)) return 0.;
    const double z13 = std::pow(Z, 1. / 3.);
    if (q >= K + m * (1. - 0.75 * SQRTE * z13)) return 0.;
    const double a[3] = {-0.00349, 148.84, -987.531};
    const double b[4] = {0.1642, 132.573, -585.361, 1407.77};
    const double c[6] = {-2.8922, -19.0156, 57.698, -63.418, 14.1166, 1.84206};
    const double d[6] = {2134.19, 581.823, -2708.85, 4767.05, 1.52918, 0.361933};
    const double energy = (K + m) * 1E+03;
    const double v = q * 1E+03 / energy;
    m *= 1E+03;
    const double Z13 = 1. / z13;
    const double rad_log = radiation_logarithm(Z);
    const double rad_log_inel = (Z == 1.) ? 446. : 1429.;
    const double Dn = 1.54 * std::pow(A, 0.27);
    const double mu_qc = m / (MMU * std::exp(1.) / Dn);
    const double rho = std::sqrt(1. + 4. * mu_qc * mu_qc);
    const double log_rho = std::log((rho + 1.) / (rho - 1.));
    const double delta1 = std::log(mu_qc) + 0.5 * rho * log_rho;
    const double delta2 = std::log(mu_qc) + 0.25 * (3. * rho - rho * rho * rho) * log_rho + 2. * mu_qc * mu_qc;
    const double delta = m * m * v / (2. * energy * (1. - v));
    double phi1 = std::log(rad_log * Z13 * (m / ME) / (1. + rad_log * Z13 * std::exp(0.5) * delta / ME));
    double phi2 = std::log(rad_log * Z13 * std::exp(-1. / 6.) * (m / ME) /
                           (1. + rad_log * Z13 * std::exp(1. / 3.) * delta / ME));
    phi1 -= delta1 * (1. - 1. / Z);
    phi2 -= delta2 * (1. - 1. / Z);
    const double s_atomic_1 = std::log(m / delta / (m * delta / (ME * ME) + SQRTE));
    const double s_atomic_2 = std::log(1. + ME / (delta * rad_log_inel * Z13 * Z13 * SQRTE));
    const double s_atomic = (4. / 3. * (1. - v) + v * v) * (s_atomic_1 - s_atomic_2);
    double s_rad;
    if (v < .0 || v > 1.) {
        s_rad = 0.;
    } else if (v < 0.02) {
        s_rad = a[0] + a[1] * v + a[2] * v * v;
    } else if (v >= 0.02 && v < 0.1) {
        s_rad = b[0] + b[1] * v + b[2] * v * v + b[3] * v * v * v;
    } else if (v >= 0.01 && v < 0.9) {
        s_rad = c[0] + c[1] * v + c[2] * v * v;
        const double tmp = std::log(1. - v);
        s_rad += c[3] * v * std::log(v) + c[4] * tmp + c[5] * tmp * tmp;
    } else {
        s_rad = d[0] + d[1] * v + d[2] * v * v;
        const double tmp = std::log(1. - v);
        s_rad += d[3] * v * std::log(v) + d[4] * tmp + d[5] * tmp * tmp;
    }
    const double result = ((2. - 2. * v + v * v) * phi1 - 2. / 3. * (1. - v) * phi2) + 1. / Z * s_atomic +
                          0.25 * ALPHA_EM * phi1 * s_rad;
    if (result <= 0.) return 0.;
    const double aux = 2 * (ME / m) * RE * Z;
    return aux * aux * (ALPHA_EM / q) * result * 1E-04;
}

// pumas.c:14294-14303
inline void pair_range(double Z, double m, double K, double* qmin, double* qmax) {
    if (qmax) {
        const double sqrte = 1.6487212707;
        const double Z13 = std::pow(Z, 1. / 3.);
        *qmax = K + m * (1. - 0.75 * sqrte * Z13);
    }
    if (qmin) *qmin = 4. * ELECTRON_MASS;
}

// pumas.c:14458-14609, doubly differential in the e+e- asymmetry rho.
inline double pair_d2_SSR(double Z, double A, double m, double K, double q, double rho) {
    const double ME = 0.5109989461, RE = 2.8179403227E-13;
    if ((Z <= 0) || (A <= 0) || (m <= 0) || (K <= 0) || (q <= 0)) return 0.;
    double qmin, qmax;
    pair_range(Z, m, K, &qmin, &qmax);
    if ((q < qmin) || (q > qmax)) return 0.;
    const double energy = (K + m) * 1E+03;
    const double v = q / (K + m);
    m *= 1E+03;
    const double rad_log = radiation_logarithm(Z);
    const double const_prefactor = 4. / (3. * PI) * Z * std::pow(ALPHA_EM * RE, 2.);
    const double Z13 = std::pow(Z, -1. / 3.);
    const double d_n = 1.54 * std::pow(A, 0.27);
    rho = 1 - rho;
    const double rho2 = rho * rho;
    double g1, g2;
    if (Z == 1.) {
        g1 = 4.4E-05;
        g2 = 4.8E-05;
    } else {
        g1 = 1.95E-05;
        g2 = 5.3E-05;
    }
    const double zeta1 = (0.073 * std::log(energy / m / (1. + g1 * std::pow(Z, 2. / 3.) * energy / m)) - 0.26);
    const double zeta2 = (0.058 * std::log(energy / m / (1 + g2 / Z13 * energy / m)) - 0.14);
    double zeta;
    if ((zeta1 > 0.) && (zeta2 > 0.))
        zeta = zeta1 / zeta2;
    else
        zeta = 0.;
    const double beta = v * v / (2. * (1. - v));
    const double xi = std::pow(m * v / (2. * ME), 2.) * (1. - rho2) / (1. - v);
    const double Be = ((2. + rho2) * (1. + beta) + xi * (3. + rho2)) * std::log(1. + 1. / xi) +
                      (1. - rho2 - beta) / (1. + xi) - (3. + rho2);
    const double Ce2 = ((1. - rho2) * (1. + beta) + xi * (3. - rho2)) * std::log(1. + 1. / xi) +
                       2. * (1. - beta - rho2) / (1. + xi) - (3. - rho2);
    const double Ce1 = Be - Ce2;
    const double De = ((2. + rho2) * (1. + beta) + xi * (3. + rho2)) * dilog(1. / (1. + xi)) -
                      (2. + rho2) * xi * std::log(1. + 1. / xi) - (xi + rho2 + beta) / (1. + xi);
    double Le1, Le2;
    if (De / Be > 0.) {
        const double tmp = De / Be;
        const double Xe = (tmp < 100) ? std::exp(-tmp) : 0.;
        Le1 = std::log(rad_log * Z13 * std::sqrt(1. + xi) /
                       (Xe + 2. * ME * std::exp(0.5) * rad_log * Z13 * (1. + xi) / (energy * v * (1. - rho2)))) -
              tmp - 0.5 * std::log(Xe + std::pow(ME / m * d_n, 2.) * (1. + xi));
        Le2 = std::log(rad_log * Z13 * std::exp(-1. / 6.) * std::sqrt(1 + xi) /
                       (Xe + 2. * ME * std::exp(1. / 3.) * rad_log * Z13 * (1. + xi) / (energy * v * (1. - rho2)))) -
              tmp - 0.5 * std::log(Xe + std::pow(ME / m * d_n, 2.) * std::exp(-1. / 3.) * (1. + xi));
    } else {
        const double tmp = De / Be;
        const double Xe_inv = (tmp > -100) ? std::exp(De / Be) : 0;
        Le1 = std::log(rad_log * Z13 * std::sqrt(1. + xi) /
                       (1. + Xe_inv * 2. * ME * std::exp(0.5) * rad_log * Z13 * (1. + xi) /
                                 (energy * v * (1. - rho2)))) -
              0.5 * De / Be - 0.5 * std::log(1. + Xe_inv * std::pow(ME / m * d_n, 2.) * (1. + xi));
        Le2 = std::log(rad_log * Z13 * std::exp(-1. / 6.) * std::sqrt(1 + xi) /
                       (1. + Xe_inv * 2. * ME * std::exp(1. / 3.) * rad_log * Z13 * (1. + xi) /
                                 (energy * v * (1. - rho2)))) -
              0.5 * De / Be - 0.5 * std::log(1. + Xe_inv * std::pow(ME / m * d_n, 2.) * std::exp(-1. / 3.) * (1. + xi));
    }
    double diagram_e = const_prefactor * (Z + zeta) * (1. - v) / v * (Ce1 * Le1 + Ce2 * Le2);
    if (diagram_e < 0.) diagram_e = 0.;
    const double Bm = ((1. + rho2) * (1. + (3. * beta) / 2) - 1. / xi * (1. + 2. * beta) * (1. - rho2)) *
                          std::log(1. + xi) +
                      xi * (1. - rho2 - beta) / (1. + xi) + (1. + 2. * beta) * (1. - rho2);
    const double Cm2 = ((1. - beta) * (1. - rho2) - xi * (1. + rho2)) * std::log(1. + xi) / xi -
                       2. * (1. - beta - rho2) / (1. + xi) + 1. - beta - (1. + beta) * rho2;
    const double Cm1 = Bm - Cm2;
    const double Dm = ((1. + rho2) * (1. + (3. * beta) / 2.) - 1. / xi * (1. + 2. * beta) * (1. - rho2)) *
                          dilog(xi / (1. + xi)) +
                      (1. + (3. * beta) / 2.) * (1. - rho2) / xi * std::log(1. + xi) +
                      (1. - rho2 - beta / 2. * (1. + rho2) + (1. - rho2) / (2. * xi) * beta) * xi / (1. + xi);
    double Lm1, Lm2;
    if (Dm / Bm > 0.) {
        const double tmp = Dm / Bm;
        const double Xm = (tmp < 100) ? std::exp(-tmp) : 0.;
        Lm1 = std::log(m / ME * rad_log * Z13 / d_n /
                       (Xm + 2. * ME * std::exp(0.5) * rad_log * Z13 * (1. + xi) / (energy * v * (1. - rho2)))) -
              tmp;
        Lm2 = std::log(m / ME * rad_log * Z13 / d_n /
                       (Xm + 2. * ME * std::exp(1. / 3.) * rad_log * Z13 * (1. + xi) / (energy * v * (1. - rho2)))) -
              tmp;
    } else {
        const double tmp = Dm / Bm;
        const double Xmv = (tmp > -100) ? std::exp(tmp) : 0.;
        Lm1 = std::log(m / ME * rad_log * Z13 / d_n /
                       (1. + 2. * ME * std::exp(0.5) * rad_log * Z13 * (1. + xi) / (energy * v * (1. - rho2)) * Xmv));
        Lm2 = std::log(m / ME * rad_log * Z13 / d_n /
                       (1. + 2. * ME * std::exp(1. / 3.) * rad_log * Z13 * (1. + xi) / (energy * v * (1. - rho2)) *
                                 Xmv));
    }
    double diagram_mu = const_prefactor * (Z + zeta) * (1. - v) / v * std::pow(ME / m, 2.) * (Cm1 * Lm1 + Cm2 * Lm2);
    if (diagram_mu < 0.) diagram_mu = 0.;
    return (diagram_e + diagram_mu) * 1E-01 / energy;
}

// pumas.c:14624-14671, atomic DCS in m^2 / GeV.
inline double pair_SSR(double Z, double A_, double mass, double K, double q) {
    if ((Z <= 0) || (A_ <= 0) || (mass <= 0) || (K <= 0) || (q <= 0)) return 0.;
    if (q <= 4. * ELECTRON_MASS) return 0.;
    const double sqrte = 1.6487212707;
    const double Z13 = std::pow(Z, 1. / 3.);
    if (q >= K + mass * (1. - 0.75 * sqrte * Z13)) return 0.;
    const double gamma = 1. + K / mass;
    const double x0 = 1. - 4. * ELECTRON_MASS / q;
    const double x1 = 1. - 6. / (gamma * (gamma - q / mass));
    double rmax;
    if ((x0 > 0) && (x1 > 0))
        rmax = std::sqrt(x0) * x1;
    else
        return 0.;
    const double ri = 1. - rmax;
    if ((ri <= 0.) || (ri >= 1.)) return 0.;
    const double tmin = std::log(ri);
    double I = 0.;
    for (int i = 0; i < 12; i++) {
        const double rho = std::exp(GQ12_X[i] * tmin);
        I -= pair_d2_SSR(Z, A_, mass, K, q, rho) * rho * GQ12_W[i] * tmin;
    }
    return (I < 0) ? 0. : I;
}

// pumas.c:14690-14745
inline double f2p_ALLM97(double x, double Q2) {
    const double m02 = 0.31985, mP2 = 49.457, mR2 = 0.15052, Q02 = 0.52544, Lambda2 = 0.06527;
    const double cP1 = 0.28067, cP2 = 0.22291, cP3 = 2.1979;
    const double aP1 = -0.0808, aP2 = -0.44812, aP3 = 1.1709;
    const double bP1 = 0.36292, bP2 = 1.8917, bP3 = 1.8439;
    const double cR1 = 0.80107, cR2 = 0.97307, cR3 = 3.4942;
    const double aR1 = 0.58400, aR2 = 0.37888, aR3 = 2.6063;
    const double bR1 = 0.01147, bR2 = 3.7582, bR3 = 0.49338;
    const double M = 0.5 * (PROTON_MASS + NEUTRON_MASS);
    const double M2 = M * M;
    const double W2 = M2 + Q2 * (1.0 / x - 1.0);
    const double t = std::log(std::log((Q2 + Q02) / Lambda2) / std::log(Q02 / Lambda2));
    const double xP = (Q2 + mP2) / (Q2 + mP2 + W2 - M2);
    const double xR = (Q2 + mR2) / (Q2 + mR2 + W2 - M2);
    const double cP = cP1 + (cP1 - cP2) * (1.0 / (1.0 + std::pow(t, cP3)) - 1.0);
    const double aP = aP1 + (aP1 - aP2) * (1.0 / (1.0 + std::pow(t, aP3)) - 1.0);
    const double bP = bP1 + bP2 * std::pow(t, bP3);
    const double cR = cR1 + cR2 * std::pow(t, cR3);
    const double aR = aR1 + aR2 * std::pow(t, aR3);
    const double bR = bR1 + bR2 * std::pow(t, bR3);
    const double F2P = cP * std::exp(aP * std::log(xP) + bP * std::log(1 - x));
    const double F2R = cR * std::exp(aR * std::log(xR) + bR * std::log(1 - x));
    return Q2 / (Q2 + m02) * (F2P + F2R);
}

// pumas.c:14752-14763
inline double shadowing_DRSS(double Z, double A, double x) {
    if (Z == 1.) return 1.;
    if (x < 0.0014)
        return std::exp(-0.1 * std::log(A));
    else if (x < 0.04)
        return std::exp((0.069 * std::log10(x) + 0.097) * std::log(A));
    else
        return 1.;
}

// pumas.c:14780-14785
inline double f2a_DRSS(double Z, double A, double F2p, double shadowing, double x) {
    return F2p * shadowing * (Z + (A - Z) * (1.0 + x * (-1.85 + x * (2.45 + x * (-2.35 + x)))));
}

// pumas.c:14987-15010
inline double photo_d2_DRSS(double Z, double A, double ml, double K, double q, double Q2) {
    const double cf = 4 * PI * ALPHA_EM * ALPHA_EM * HBAR_C * HBAR_C;
    const double M = 0.5 * (PROTON_MASS + NEUTRON_MASS);
    const double E = K + ml;
    const double y = q / E;
    const double x = 0.5 * Q2 / (M * q);
    const double F2p = f2p_ALLM97(x, Q2);
    const double shadowing = shadowing_DRSS(Z, A, x);
    const double F2A = f2a_DRSS(Z, A, F2p, shadowing, x);
    const double R = 0.;
    const double dds = (1 - y + 0.5 * (1 - 2 * ml * ml / Q2) * (y * y + Q2 / (E * E)) / (1 + R)) / (Q2 * Q2) -
                       0.25 / (E * E * Q2);
    return cf * F2A * dds / q;
}

// pumas.c:15066-15120 (mode 0) with the DRSS DDCS (pumas.c:15142-15146);
// the log-scale 9-point rule of math_gauss_quad_initialise (pumas.c:12627-12660).
inline double photo_DRSS(double Z, double A, double ml, double K, double q) {
    if ((Z <= 0) || (A <= 0) || (ml <= 0) || (K <= 0)) return 0.;
    const double M = 0.5 * (NEUTRON_MASS + PROTON_MASS);
    const double mpi = PION_MASS;
    if ((q <= mpi + 0.5 * mpi * mpi / M) || (q >= K + ml - 0.5 * (M + ml * ml / M))) return 0.;
    const double E = K + ml;
    const double ml2 = ml * ml;
    const double Q2min = ml2 * (q * q - 0.5 * ml2) / (E * (E - q));
    const double Q2max = 2.0 * M * (q - mpi) - mpi * mpi;
    if ((Q2max < Q2min) | (Q2min < 0)) return 0.;
    const double lmin = std::log(Q2min);
    const double lmax = std::log(Q2max);
    const double dx = lmax - lmin;
    double ds = 0.;
    for (int i = 0; i < 9; i++) {
        const double Q2 = std::exp(lmin + dx * GQ9_X[i]);
        const double w = GQ9_W[i] * (dx * Q2);
        const double tmp = photo_d2_DRSS(Z, A, ml, K, q, Q2);
        if (tmp <= 0.) continue;
        ds += tmp * w;
    }
    return (ds < 0.) ? 0. : ds;
}

// pumas.c:11114-11130
inline void photo_range(double m, double K, double* qmin, double* qmax) {
    const double M = 0.5 * (NEUTRON_MASS + PROTON_MASS);
    const double mpi = PION_MASS;
    if (qmin) *qmin = mpi + 0.5 * mpi * mpi / M;
    if (qmax) *qmax = K + m - 0.5 * (M + m * m / M);
}

// The m^2/kg wrappers dcs_bremsstrahlung / dcs_pair_production /
// dcs_photonuclear of pumas.c:11067-11112 (processes 0, 1, 2).
inline double dcs(int process, double Z, double A, double mass, double K, double q) {
    if (process == 0) return brems_SSR(Z, A, mass, K, q) * 1E+03 * AVOGADRO_NUMBER * (mass + K) / A;
    if (process == 1) return pair_SSR(Z, A, mass, K, q) * 1E+03 * AVOGADRO_NUMBER * (mass + K) / A;
    double qmin, qmax;
    photo_range(mass, K, &qmin, &qmax);
    if ((q < qmin) || (q > qmax)) return 0.;  // dcs_photonuclear_check, pumas.c:11138-11143
    return photo_DRSS(Z, A, mass, K, q) * 1E+03 * AVOGADRO_NUMBER * (mass + K) / A;
}

// pumas.c:1548-1554
inline int n_table_dcs(double cutoff) {
    int n_table = DCS_MODEL_N_REVERSE;
    if (cutoff < DCS_MODEL_X_REVERSE) {
        int n = (int)(std::log10(DCS_MODEL_X_REVERSE / cutoff) * DCS_MODEL_PER_DECADE);
        if (n < DCS_MODEL_N_MIN) n = DCS_MODEL_N_MIN;
        n_table += n;
    }
    return n_table;
}

// The x = ln(q / K) grid of compute_dcs_table (pumas.c:11712-11737).
inline void x_grid(double cutoff, int n, double* x) {
    const int p = DCS_MODEL_N_REVERSE;
    const int m = n - p;
    double xrev;
    if (m > 0) {
        xrev = DCS_MODEL_X_REVERSE;
        const double x0 = std::log(cutoff);
        const double dx = std::log(xrev / cutoff) / (m - 1);
        for (int i = 0; i < m; i++) x[i] = x0 + i * dx;
    } else {
        xrev = cutoff;
    }
    const double dxmin = DCS_MODEL_DXMIN;
    const double dx = std::log((DCS_MODEL_MAX_FRACTION - xrev) / dxmin) / ((m > 0) ? p : p - 1);
    for (int i = 0; i < p; i++) x[n - 1 - i] = std::log(DCS_MODEL_MAX_FRACTION - dxmin * std::exp(i * dx));
}

// pumas.c:10010-10020
inline double diff3(double x0, double x1, double x2, double y0, double y1, double y2) {
    const double h1 = x1 - x0;
    const double h2 = x2 - x0;
    const double delta = h1 * h2 * (h2 - h1);
    const double c1 = h2 * h2 / delta;
    const double c2 = -h1 * h1 / delta;
    const double c0 = -(c1 + c2);
    return c0 * y0 + c1 * y1 + c2 * y2;
}

// pumas.c:10039-10088 with der = 0.
inline void pchip_initialise(int n, const double* x, const double* y, double* m) {
    if (n >= 3) {
        m[0] = (y[1] - y[0]) / (x[1] - x[0]);
        n--;
        x++;
        y++;
        m++;
    }
    if (n == 1) {
        m[0] = 0;
    } else if (n == 2) {
        const double d = (y[1] - y[0]) / (x[1] - x[0]);
        m[0] = m[1] = d;
    } else {
        for (int i = 1; i < n - 1; i++) {
            const double h1 = x[i] - x[i - 1];
            const double h2 = x[i + 1] - x[i];
            const double S1 = (y[i] - y[i - 1]) / h1;
            const double S2 = (y[i + 1] - y[i]) / h2;
            const double tmp = S1 * S2;
            if (tmp > 0) {
                const double a = (h1 + 2 * h2) / (3 * (h1 + h2));
                m[i] = tmp / ((1 - a) * S1 + a * S2);
            } else {
                m[i] = 0.;
            }
        }
        m[0] = diff3(x[0], x[1], x[2], y[0], y[1], y[2]);
        m[n - 1] = diff3(x[n - 1], x[n - 2], x[n - 3], y[n - 1], y[n - 2], y[n - 3]);
    }
}

// One row of the table (dcs_tabulate_row, pumas.c:11422-11507): data[2 i] =
// ln DCS, data[2 i + 1] = PCHIP slope at node i over the first run of
// positive DCS values, zeros elsewhere; *alpha the envelope exponent. The
// work arrays y and m hold n doubles each; data_d, when given, receives the
// same row in FP64 (before the float rounding of the table). (The caller's
// cross-section normalisation gate, which zeroes a whole row, is the
// caller's.)
inline void tabulate_row(int process, double Z, double A, double mass, double cutoff, double K, int n,
                         const double* x, double* y, double* m, float* data, float* alpha,
                         double* data_d = nullptr) {
    int imin = n, imax = n - 1;
    for (int i = 0; i < n; i++) {
        const double q = K * std::exp(x[i]);
        const double d = dcs(process, Z, A, mass, K, q);
        if (d > 0) {
            if (i < imin) imin = i;
            y[i] = std::log(d);
        } else if (imin < n) {
            imax = i - 1;
            break;
        }
    }
    if (imin <= imax) pchip_initialise(imax - imin + 1, x + imin, y + imin, m + imin);
    double qmin, qmax;
    if (process == 2) {
        photo_range(mass, K, &qmin, &qmax);
        qmin += qmin;
    } else {
        pair_range(Z, mass, K, &qmin, &qmax);
        if (process == 0)
            qmin = 0.5 * cutoff * K;
        else
            qmin *= 4;
    }
    const double qrev = DCS_MODEL_X_REVERSE * K;
    qmax *= 0.5;
    if (qmax > qrev) qmax = qrev;
    const double xmin = std::log(qmin / K), xmax = std::log(qmax / K);
    int ia0 = imax, ia1 = imin;
    for (int i = 0; i < n; i++) {
        if (data_d) {
            const bool in = (i >= imin) && (i <= imax);
            data_d[2 * i] = in ? y[i] : 0.;
            data_d[2 * i + 1] = in ? m[i] : 0.;
        }
        if ((i >= imin) && (i <= imax)) {
            data[2 * i] = (float)y[i];
            data[2 * i + 1] = (float)m[i];
            if ((x[i] > xmin) && (x[i] < xmax)) {
                if (ia0 > i) ia0 = i;
                if (ia1 < i) ia1 = i;
            }
        } else {
            data[2 * i] = 0.f;
            data[2 * i + 1] = 0.f;
        }
    }
    if ((ia0 < ia1) && (ia0 < n - 1)) {
        const int ia = (3 * ia0 + ia1) / 4;
        double a = -m[ia];
        if (std::fabs(a - 1.) <= 0.1) a = 1.;
        *alpha = (float)a;
    } else {
        *alpha = 0.f;
    }
}

// compute_dcs_table's low-energy fill (pumas.c:11745-11761): rows below the
// first one with a positive exponent take that exponent. alpha[row] over
// n_K rows of one process and element.
inline void fill_low_energy_alpha(int n_K, float* alpha, int stride) {
    int n0 = 0;
    float a = 0.f;
    for (int row = 0; row < n_K; row++) {
        if (alpha[row * stride] > 0.f) {
            n0 = row;
            a = alpha[row * stride];
            break;
        }
    }
    for (int row = 0; row < n0; row++) alpha[row * stride] = a;
}

// ---- the envelope maximum (dcs_tabulate_envelope_row, pumas.c:11509-11679) ----

// One process and element of the table as the envelope search reads it: the
// rows [n_K][2 n] (float), the nodes xt (float) and the energy grid. The
// reference reads the nodes from physics->table_DCS_x, which
// compute_dcs_table writes only after element 0 is done (pumas.c:11775-11783),
// so the search of element 0 sees zeros there (and takes the model
// everywhere): callers pass a zero node array for element 0.
struct EnvTable {
    int process;
    double Z, A, mass;
    const double* K;
    int n_K;
    const float* rows;
    const float* xt;
    int n;
};

// table_bracketf (pumas.c:4311-4319), as a loop.
inline void bracketf(const float* table, float value, int* p1, int* p2) {
    while (true) {
        const int i3 = (*p1 + *p2) / 2;
        if (value >= table[i3])
            *p1 = i3;
        else
            *p2 = i3;
        if (*p2 - *p1 < 2) break;
    }
}

// table_index without a context (pumas.c:4249-4290): the row of K in the
// energy grid, -1 below it, n_K - 1 at or above its last node.
inline int table_index(const double* table, int n_K, double value) {
    const int imax = n_K - 1;
    if (value < table[0]) return -1;
    if (value >= table[imax]) return imax;
    int i1 = 0, i2 = imax;
    while (true) {  // table_bracket (pumas.c:4299-4306)
        const int i3 = (i1 + i2) / 2;
        if (value >= table[i3])
            i1 = i3;
        else
            i2 = i3;
        if (i2 - i1 < 2) break;
    }
    return i1;
}

// dcs_evaluate without a simulation context (pumas.c:11307-11398) for
// processes 0-2: the table's cubic Hermite spline in x = ln(q / K), linear in
// ln K between rows, times the Jacobian K / (K + m); the model itself below
// 2 q_min (photonuclear) or 4 q_min (pair production), outside the nodes'
// range, and where one of the four bracketing nodes is empty. (An energy
// below the grid, where the reference reads before the table, takes the
// model here; the envelope search never asks for one.)
inline double dcs_interpolated(const EnvTable& tb, double K, double q) {
    const double wj = K / (K + tb.mass);
    double qlow = 0., qmin, qmax;
    if (tb.process == 2) {
        photo_range(tb.mass, K, &qmin, &qmax);
        if ((q < qmin) || (q > qmax)) return 0.;
        qlow = 2 * qmin;
    } else {
        pair_range(tb.Z, tb.mass, K, &qmin, &qmax);
        if (tb.process == 1)
            qlow = 4 * qmin;
        else
            qmin = 0.;
        if ((q < qmin) || (q > qmax)) return 0.;
    }
    if (q < qlow) return dcs(tb.process, tb.Z, tb.A, tb.mass, K, q) * wj;
    const int nx = tb.n;
    const float xmin = tb.xt[0], xmax = tb.xt[nx - 1];
    const float x = logf((float)(q / K));
    if ((x < xmin) || (x >= xmax)) return dcs(tb.process, tb.Z, tb.A, tb.mass, K, q) * wj;
    int ix = 0, tmpi = nx - 1;
    bracketf(tb.xt, x, &ix, &tmpi);
    if ((ix < 0) || (ix >= nx - 1)) return dcs(tb.process, tb.Z, tb.A, tb.mass, K, q) * wj;
    float p00, p01, m00, m01, p10, p11, m10, m11, h0, h1;
    const int imax = tb.n_K - 1;
    const double Kmax = tb.K[imax];
    int i0 = 0;
    if ((K >= Kmax) || ((i0 = table_index(tb.K, tb.n_K, K)) >= imax)) {
        const float* data = tb.rows + int64_t(imax) * 2 * nx;
        p00 = p10 = data[2 * ix];
        m00 = m10 = data[2 * ix + 1];
        p01 = p11 = data[2 * ix + 2];
        m01 = m11 = data[2 * ix + 3];
        h0 = 1.f;
        h1 = 0.f;
    } else {
        if (i0 < 0) return dcs(tb.process, tb.Z, tb.A, tb.mass, K, q) * wj;
        const float* data = tb.rows + int64_t(i0) * 2 * nx;
        p00 = data[2 * ix];
        m00 = data[2 * ix + 1];
        p01 = data[2 * ix + 2];
        m01 = data[2 * ix + 3];
        data += 2 * nx;
        p10 = data[2 * ix];
        m10 = data[2 * ix + 1];
        p11 = data[2 * ix + 2];
        m11 = data[2 * ix + 3];
        const float K0 = (float)tb.K[i0];
        const float K1 = (float)tb.K[i0 + 1];
        h1 = logf(((float)K) / K0) / logf(K1 / K0);
        h0 = 1.f - h1;
    }
    if (((p00 == 0.f) && (m00 == 0.f)) || ((p01 == 0.f) && (m01 == 0.f)) || ((p10 == 0.f) && (m10 == 0.f)) ||
        ((p11 == 0.f) && (m11 == 0.f)))
        return dcs(tb.process, tb.Z, tb.A, tb.mass, K, q) * wj;
    const float x0 = tb.xt[ix];
    const float x1 = tb.xt[ix + 1];
    const float dx = x1 - x0;
    m00 *= dx;
    m01 *= dx;
    const float t = (x - x0) / dx;
    const float c02 = -3 * (p00 - p01) - 2 * m00 - m01;
    const float c03 = 2 * (p00 - p01) + m00 + m01;
    float r = expf(p00 + t * (m00 + t * (c02 + t * c03))) * h0;
    if (h1 > 0.f) {
        m10 *= dx;
        m11 *= dx;
        const float c12 = -3 * (p10 - p11) - 2 * m10 - m11;
        const float c13 = 2 * (p10 - p11) + m10 + m11;
        r += expf(p10 + t * (m10 + t * (c12 + t * c13))) * h1;
    }
    return r * wj;
}

// dcs_tabulate_envelope_objective (pumas.c:11516-11524): -DCS x^alpha at ln x.
struct EnvObjective {
    const EnvTable& tb;
    double K, alpha;
    double operator()(double lnx) const {
        const double x = std::exp(lnx);
        const double q = x * K;
        return -dcs_interpolated(tb, K, q) * std::pow(x, alpha);
    }
};

// math_find_minimum_bisection (pumas.c:12416-12459).
template <class F>
inline int find_minimum_bisection(const F& f, double xa, double xb, double xc, const double* fb_p, double tol,
                                  int max_iter, double* fopt) {
    double fb = fb_p ? *fb_p : f(xb);
    tol *= std::fabs(xc - xa);
    int i = 0;
    while ((xc - xa > tol) && (i < max_iter)) {
        if (xb - xa > xc - xb) {
            const double x = 0.5 * (xa + xb);
            const double fi = f(x);
            if (fi < fb) {
                xc = xb;
                xb = x;
                fb = fi;
            } else {
                xa = x;
            }
        } else {
            const double x = 0.5 * (xb + xc);
            const double fi = f(x);
            if (fi < fb) {
                xa = xb;
                xb = x;
                fb = fi;
            } else {
                xc = x;
            }
        }
        i++;
    }
    *fopt = fb;
    return (i <= max_iter) ? i : -1;
}

// math_find_minimum with algo 0 (pumas.c:12172-12232): a bisection that
// brackets a minimum, then the three-point bisection.
template <class F>
inline int find_minimum(const F& f, double xa, double xc, const double* fa_p, const double* fc_p, double tol,
                        int max_iter, double* fopt) {
    double fa = fa_p ? *fa_p : f(xa);
    double fc = fc_p ? *fc_p : f(xc);
    const double atol = tol * std::fabs(xc - xa);
    double xb = 0, fb = 0;
    int i;
    for (i = 0; i < max_iter; i++) {
        xb = 0.5 * (xa + xc);
        fb = f(xb);
        if ((fb < fa) && (fb < fc)) break;
        if (std::fabs(xc - xa) <= atol) {
            *fopt = fb;
            return i;
        }
        if (fb < fc) {
            xc = xb;
            fc = fb;
        } else {
            xa = xb;
            fa = fb;
        }
    }
    if (i == max_iter) {
        if (fa < fb) {
            fb = fa;
            xb = xa;
        } else if (fc < fb) {
            fb = fc;
            xb = xc;
        }
        *fopt = fb;
        return -1;
    }
    const int status = find_minimum_bisection(f, xa, xb, xc, &fb, tol, max_iter - i, fopt);
    return (status < 0) ? -2 : status + i;
}

// dcs_tabulate_envelope_row (pumas.c:11526-11679): the maximum of DCS x^alpha
// over the row's kinematic range above the cutoff, for the row's envelope
// exponent alpha; 0 where the row's cross-section normalisation cs is not
// positive or the range is empty.
inline float envelope_max(const EnvTable& tb, int row, double cutoff, float alpha, double cs) {
    constexpr double RESOLUTION = 1E-10;
    if (cs <= 0.) return 0.f;
    const double K = tb.K[row];
    double qmin, qmax;
    if (tb.process == 2) {  // dcs_get_range (pumas.c:11047-11056)
        photo_range(tb.mass, K, &qmin, &qmax);
    } else {
        pair_range(tb.Z, tb.mass, K, &qmin, &qmax);
        if (tb.process == 0) qmin = 0.;
    }
    double xmin = qmin / K;
    if (xmin < cutoff) xmin = cutoff;
    xmin = std::log(xmin);
    double xmax = std::log(qmax / K);
    if (xmin >= xmax) return 0.f;
    const EnvObjective f{tb, K, (double)alpha};
    double f0 = f(xmin);
    double f1 = f(xmax);
    if (f0 == 0.) {
        double x0 = xmin, x1 = xmax;
        while (x1 - x0 > RESOLUTION) {
            const double x2 = 0.5 * (x0 + x1);
            const double f2 = f(x2);
            if (f2 != 0) {
                x1 = xmin = x2;
                f0 = f2;
            } else {
                x0 = x2;
            }
        }
    }
    if (f1 == 0.) {
        double x0, x1;
        if (K >= 1E+01) {  // an exponential decrement from the end
            const double xm = std::exp(xmax);
            double delta = RESOLUTION, x = xmax, fx = f1;
            while (xm > delta) {
                x = std::log(xm - delta);
                fx = f(x);
                if (fx != 0.)
                    break;
                else
                    delta += delta;
            }
            f1 = fx;
            x0 = xmax = x;
            x1 = std::log(xm - delta / 2);
        } else {
            x0 = xmin;
            x1 = xmax;
        }
        while (x1 - x0 > RESOLUTION) {
            const double x2 = 0.5 * (x0 + x1);
            const double f2 = f(x2);
            if (f2 != 0) {
                x0 = xmax = x2;
                f1 = f2;
            } else {
                x1 = x2;
            }
        }
    }
    const double xm = 0.5 * (xmin + xmax);
    const double fm = f(xm);
    double fopt0 = 0., fopt1 = 0.;
    find_minimum(f, xmin, xm, &f0, &fm, RESOLUTION, 100, &fopt0);
    find_minimum(f, xm, xmax, &fm, &f1, RESOLUTION, 100, &fopt1);
    if ((tb.process == 1) && (K >= 1E+07)) {
        // a sharp peak next to xmax: a logarithmic scan for a bracketing
        constexpr int N_SCAN = 7;
        double xs[N_SCAN], fs[N_SCAN], fmin = 0.;
        int imin = -1;
        for (int i = 0; i < N_SCAN; i++) {
            xs[i] = xmax + std::log(1. - std::pow(10, -(i + 1)));
            fs[i] = f(xs[i]);
            if (fs[i] < fmin) {
                imin = i;
                fmin = fs[i];
            }
        }
        if ((imin > 0) && (imin < N_SCAN - 1)) {
            double fopt2 = 0.;
            find_minimum_bisection(f, xs[imin - 1], xs[imin], xs[imin + 1], fs + imin, RESOLUTION, 100, &fopt2);
            if (fopt2 < fopt1) fopt1 = fopt2;
        }
    }
    double fopt = (fopt0 < fopt1) ? fopt0 : fopt1;
    if (f0 < fopt) fopt = f0;
    if (f1 < fopt) fopt = f1;
    return (float)(-fopt);
}

}  // namespace pumas
}  // namespace noa_oracle
