// PUMAS v1.2.1 DCS tabulation on the GPU: the x = ln(q / K) grid and the
// per-row table of compute_dcs_table (src/noa/3rdparty/_pumas/pumas.c:
// 11693-11787) with its row fill dcs_tabulate_row (:11422-11507), for the
// default models PUMAS tabulates (pumas.c:190-198): SSR bremsstrahlung
// (:13979-14068), SSR pair production (:14458-14671, 12-point rule over the
// asymmetry) and DRSS photonuclear with the ALLM97 structure function
// (:14690-15146, 9-point log-scale rule over Q^2), through their m^2/kg
// wrappers (:11067-11112).
//
// One wave per table row (process, element, kinetic energy): each lane
// evaluates the DCS at its grid nodes (the 12- and 9-point integrals run
// inside the lane), the row's logarithms and abscissae go to LDS, and the
// lanes then take the PCHIP slopes (Fritsch-Butland interior, three-point
// ends and the linear first bin of math_pchip_initialise, pumas.c:10039-10088)
// and the envelope exponent. A second launch gives the rows below the first
// positive exponent of each process and element that exponent
// (pumas.c:11745-11761), and a third the envelope maxima (pumas.c:11526-11679,
// a search over the interpolated table; see pumas_envelope_kernel).
//
// FP64 throughout, with the device libm (log, exp, pow, sqrt, log10); the
// grid is computed on the host with the reference's operations.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "capi_util.hh"

namespace noa_hip {
namespace pumas {

constexpr int MAX_X = 192;  // grid nodes: cutoff >= ~2e-12
constexpr int N_REVERSE = 20;
constexpr double X_REVERSE = 0.6;
constexpr double DXMIN = 1E-03;
constexpr double MAX_FRACTION = 0.95;
constexpr int N_MIN = 10;
constexpr int PER_DECADE = 15;

constexpr double ALPHA_EM = 7.2973525693E-03;
constexpr double HBAR_C = 1.973269804E-16;
constexpr double ELECTRON_MASS = 0.510998910E-03;
constexpr double PROTON_MASS = 0.938272;
constexpr double NEUTRON_MASS = 0.939565;
constexpr double PION_MASS = 0.13957018;
constexpr double AVOGADRO_NUMBER = 6.02214076E+23;
constexpr double PI = 3.14159265358979323846;
constexpr double NUCLEON_MASS = 0.5 * (PROTON_MASS + NEUTRON_MASS);
// MeV-unit constants of the SSR models (pumas.c:13891-13894, 14461-14463)
constexpr double SSR_SQRTE = 1.648721270700128;
constexpr double SSR_ME = 0.5109989461;
constexpr double SSR_RE = 2.8179403227E-13;
constexpr double SSR_MMU = 105.6583745;
constexpr double PAIR_SQRTE = 1.6487212707;  // pumas.c:14298, 14633

// Gauss-Legendre nodes and weights on [0, 1] (the 9- and 12-point rows of
// pumas.c:12527-12614).
static __constant__ double GQ9_X[9] = {0.0159198802461870, 0.0819844463366821, 0.1933142836497048,
                                       0.3378732882980955, 0.5000000000000000, 0.6621267117019045,
                                       0.8066857163502952, 0.9180155536633179, 0.9840801197538130};
static __constant__ double GQ9_W[9] = {0.0406371941807872, 0.0903240803474287, 0.1303053482014677,
                                       0.1561735385200015, 0.1651196775006299, 0.1561735385200015,
                                       0.1303053482014677, 0.0903240803474287, 0.0406371941807872};
static __constant__ double GQ12_X[12] = {0.0092196828766404, 0.0479413718147625, 0.1150486629028477,
                                         0.2063410228566913, 0.3160842505009099, 0.4373832957442655,
                                         0.5626167042557344, 0.6839157494990901, 0.7936589771433087,
                                         0.8849513370971523, 0.9520586281852375, 0.9907803171233596};
static __constant__ double GQ12_W[12] = {0.0235876681932559, 0.0534696629976592, 0.0800391642716731,
                                         0.1015837133615330, 0.1167462682691774, 0.1245735229067014,
                                         0.1245735229067014, 0.1167462682691774, 0.1015837133615330,
                                         0.0800391642716731, 0.0534696629976592, 0.0235876681932559};

// Chebyshev coefficients of the real dilogarithm (CERNLIB RDILOG, C332), the
// series of pumas.c:12673-12738.
static __constant__ double DILOG_C[20] = {
    0.42996693560813697,  0.40975987533077105,  -0.01858843665014592, 0.00145751084062268,
    -0.00014304184442340, 0.00001588415541880,  -0.00000190784959387, 0.00000024195180854,
    -0.00000003193341274, 0.00000000434545063,  -0.00000000060578480, 0.00000000008612098,
    -0.00000000001244332, 0.00000000000182256,  -0.00000000000027007, 0.00000000000004042,
    -0.00000000000000610, 0.00000000000000093,  -0.00000000000000014, 0.00000000000000002};

struct Grid {
    double x[MAX_X];
    int32_t n;
    double cutoff;
};

// Launch invariants computed on the host with the host libm, the
// reference's: the element's powers and logarithms, the mass-dependent
// bremsstrahlung terms and the literal exponentials, so that every exit
// (q >= q_max, ...) is decided on the reference's values.
struct Globals {
    double e1, e05, e_16, e13, e_13;  // exp(1), exp(1/2), exp(-1/6), exp(1/3), exp(-1/3)
    double lq0;                       // log(Q0^2 / Lambda^2) of ALLM97
};

struct ElementConsts {
    double Z, A;
    double rl;          // radiation_logarithm(Z)
    double z13;         // Z^(1/3)
    double Z13m;        // Z^(-1/3)
    double Z23;         // Z^(2/3)
    double dn;          // 1.54 A^0.27
    double lnA;         // ln A
    double shadow_low;  // exp(-0.1 ln A)
    double b_qoff;      // m (1 - 0.75 sqrt(e) Z^(1/3)), bremsstrahlung's sqrt(e)
    double p_qoff;      // the same with pair production's sqrt(e)
    double b_delta1;    // the bremsstrahlung nuclear-size terms (pumas.c:14009-14016)
    double b_delta2;
};

constexpr int MAX_EL = 16;  // elements per table launch (kernel-argument budget)

struct ElementBlock {
    ElementConsts e[MAX_EL];
};

// Li2(x) by the RDILOG reductions onto [0, 1] and the Clenshaw sum.
__device__ double dilog(double x) {
#pragma clang fp contract(off)
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
        const double l1 = log(-T), l2 = log(1. + 1. / T);
        A = -PI3 + 0.5 * (l1 * l1 - l2 * l2);
    } else if (T < -1.) {
        Y = -1. - T;
        S = -1.;
        A = log(-T);
        A = -PI6 + A * (A + log(1. + 1. / T));
    } else if (T <= -0.5) {
        Y = -(1. + T) / T;
        S = 1.;
        A = log(-T);
        A = -PI6 + A * (-0.5 * A + log(1. + T));
    } else if (T < 0.) {
        Y = -T / (1. + T);
        S = -1.;
        const double l = log(1 + T);
        A = 0.5 * (l * l);
    } else if (T <= 1.) {
        Y = T;
        S = 1.;
        A = 0.;
    } else {
        Y = 1. / T;
        S = -1.;
        const double l = log(T);
        A = PI6 + 0.5 * (l * l);
    }
    const double H = Y + Y - 1;
    const double ALFA = H + H;
    double B0 = 0., B1 = 0., B2 = 0.;
#pragma unroll
    for (int i = 19; i >= 0; i--) {
        B0 = DILOG_C[i] + ALFA * B1 - B2;
        B2 = B1;
        B1 = B0;
    }
    return -(S * (B0 - H * B2) + A);
}

// SSR bremsstrahlung, atomic DCS in m^2 / GeV (pumas.c:13979-14068).
__device__ double brems_ssr(const ElementConsts& c, const Globals& G, double m, double K, double q) {
#pragma clang fp contract(off)
    const double Z = c.Z;
    if ((Z <= 0) || (c.A <= 0) || (m <= 0) || (K <= 0) || (q <= 0)) return 0.;
    if (q >= K + c.b_qoff) return 0.;
    const double energy = (K + m) * 1E+03;
    const double v = q * 1E+03 / energy;
    m *= 1E+03;
    const double Z13 = 1. / c.z13;
    const double rl_inel = (Z == 1.) ? 446. : 1429.;
    const double delta = m * m * v / (2. * energy * (1. - v));
    const double rz = c.rl * Z13;
    double phi1 = log(rz * (m / SSR_ME) / (1. + rz * G.e05 * delta / SSR_ME));
    double phi2 = log(rz * G.e_16 * (m / SSR_ME) / (1. + rz * G.e13 * delta / SSR_ME));
    const double iz = 1. - 1. / Z;
    phi1 -= c.b_delta1 * iz;
    phi2 -= c.b_delta2 * iz;
    const double s1 = log(m / delta / (m * delta / (SSR_ME * SSR_ME) + SSR_SQRTE));
    const double s2 = log(1. + SSR_ME / (delta * rl_inel * Z13 * Z13 * SSR_SQRTE));
    const double s_atomic = (4. / 3. * (1. - v) + v * v) * (s1 - s2);
    double s_rad;
    if (v < .0 || v > 1.) {
        s_rad = 0.;
    } else if (v < 0.02) {
        s_rad = -0.00349 + 148.84 * v + -987.531 * v * v;
    } else if (v < 0.1) {
        s_rad = 0.1642 + 132.573 * v + -585.361 * v * v + 1407.77 * v * v * v;
    } else if (v < 0.9) {
        s_rad = -2.8922 + -19.0156 * v + 57.698 * v * v;
        const double t = log(1. - v);
        s_rad += -63.418 * v * log(v) + 14.1166 * t + 1.84206 * t * t;
    } else {
        s_rad = 2134.19 + 581.823 * v + -2708.85 * v * v;
        const double t = log(1. - v);
        s_rad += 4767.05 * v * log(v) + 1.52918 * t + 0.361933 * t * t;
    }
    const double result =
        ((2. - 2. * v + v * v) * phi1 - 2. / 3. * (1. - v) * phi2) + 1. / Z * s_atomic + 0.25 * ALPHA_EM * phi1 * s_rad;
    if (result <= 0.) return 0.;
    const double aux = 2 * (SSR_ME / m) * SSR_RE * Z;
    return aux * aux * (ALPHA_EM / q) * result * 1E-04;
}

// The parts of the SSR pair DDCS (pumas.c:14466-14500) that do not depend
// on the asymmetry rho, for one (element, K, q).
struct PairQ {
    double v, m, energy, ev, rl, Z13, d_n, beta, xi0, pref, me_dn2, me_m2;
};

__device__ __forceinline__ PairQ pair_setup(const ElementConsts& c, double m_gev, double K, double q) {
#pragma clang fp contract(off)
    const double Z = c.Z;
    PairQ s;
    s.energy = (K + m_gev) * 1E+03;
    s.v = q / (K + m_gev);
    s.m = m_gev * 1E+03;
    s.ev = s.energy * s.v;
    s.rl = c.rl;
    const double are = ALPHA_EM * SSR_RE;
    const double const_prefactor = 4. / (3. * PI) * Z * (are * are);
    s.Z13 = c.Z13m;
    s.d_n = c.dn;
    double g1, g2;
    if (Z == 1.) {
        g1 = 4.4E-05;
        g2 = 4.8E-05;
    } else {
        g1 = 1.95E-05;
        g2 = 5.3E-05;
    }
    const double zeta1 = (0.073 * log(s.energy / s.m / (1. + g1 * c.Z23 * s.energy / s.m)) - 0.26);
    const double zeta2 = (0.058 * log(s.energy / s.m / (1 + g2 / s.Z13 * s.energy / s.m)) - 0.14);
    const double zeta = ((zeta1 > 0.) && (zeta2 > 0.)) ? zeta1 / zeta2 : 0.;
    s.pref = const_prefactor * (Z + zeta) * (1. - s.v) / s.v;
    s.beta = s.v * s.v / (2. * (1. - s.v));
    const double mv = s.m * s.v / (2. * SSR_ME);
    s.xi0 = mv * mv;
    const double a = SSR_ME / s.m * s.d_n;
    s.me_dn2 = a * a;
    const double b = SSR_ME / s.m;
    s.me_m2 = b * b;
    return s;
}

// SSR pair DDCS at the reference's rho (pumas.c:14499-14609).
__device__ double pair_d2_ssr(const PairQ& s, const Globals& G, double rho) {
#pragma clang fp contract(off)
    rho = 1 - rho;
    const double rho2 = rho * rho;
    const double beta = s.beta;
    const double omr = 1. - rho2;
    const double xi = s.xi0 * omr / (1. - s.v);
    const double xi1 = 1. + xi;
    const double w = s.ev * omr;
    const double c05 = 2. * SSR_ME * G.e05 * s.rl * s.Z13 * xi1 / w;
    const double c13 = 2. * SSR_ME * G.e13 * s.rl * s.Z13 * xi1 / w;

    // electron diagram
    const double lxi_i = log(1. + 1. / xi);
    const double Pe = (2. + rho2) * (1. + beta) + xi * (3. + rho2);
    const double Be = Pe * lxi_i + (1. - rho2 - beta) / xi1 - (3. + rho2);
    const double Ce2 = (omr * (1. + beta) + xi * (3. - rho2)) * lxi_i + 2. * (1. - beta - rho2) / xi1 - (3. - rho2);
    const double Ce1 = Be - Ce2;
    const double De = Pe * dilog(1. / xi1) - (2. + rho2) * xi * lxi_i - (xi + rho2 + beta) / xi1;
    const double sx = s.rl * s.Z13 * sqrt(xi1);
    const double sx6 = s.rl * s.Z13 * G.e_16 * sqrt(xi1);
    const double re = De / Be;
    double Le1, Le2;
    if (re > 0.) {
        const double Xe = (re < 100) ? exp(-re) : 0.;
        Le1 = log(sx / (Xe + c05)) - re - 0.5 * log(Xe + s.me_dn2 * xi1);
        Le2 = log(sx6 / (Xe + c13)) - re - 0.5 * log(Xe + s.me_dn2 * G.e_13 * xi1);
    } else {
        const double Xi = (re > -100) ? exp(re) : 0.;
        Le1 = log(sx / (1. + Xi * c05)) - 0.5 * re - 0.5 * log(1. + Xi * s.me_dn2 * xi1);
        Le2 = log(sx6 / (1. + Xi * c13)) - 0.5 * re - 0.5 * log(1. + Xi * s.me_dn2 * G.e_13 * xi1);
    }
    double diagram_e = s.pref * (Ce1 * Le1 + Ce2 * Le2);
    if (diagram_e < 0.) diagram_e = 0.;

    // muon diagram
    const double lxi = log(xi1);
    const double b15 = 1. + (3. * beta) / 2.;
    const double Pm = (1. + rho2) * b15 - 1. / xi * (1. + 2. * beta) * omr;
    const double Bm = Pm * lxi + xi * (1. - rho2 - beta) / xi1 + (1. + 2. * beta) * omr;
    const double Cm2 = ((1. - beta) * omr - xi * (1. + rho2)) * lxi / xi - 2. * (1. - beta - rho2) / xi1 + 1. - beta -
                       (1. + beta) * rho2;
    const double Cm1 = Bm - Cm2;
    const double Dm = Pm * dilog(xi / xi1) + b15 * omr / xi * lxi +
                      (1. - rho2 - beta / 2. * (1. + rho2) + omr / (2. * xi) * beta) * xi / xi1;
    const double num = s.m / SSR_ME * s.rl * s.Z13 / s.d_n;
    const double rm = Dm / Bm;
    double Lm1, Lm2;
    if (rm > 0.) {
        const double Xm = (rm < 100) ? exp(-rm) : 0.;
        Lm1 = log(num / (Xm + c05)) - rm;
        Lm2 = log(num / (Xm + c13)) - rm;
    } else {
        const double Xv = (rm > -100) ? exp(rm) : 0.;
        Lm1 = log(num / (1. + c05 * Xv));
        Lm2 = log(num / (1. + c13 * Xv));
    }
    double diagram_mu = s.pref * s.me_m2 * (Cm1 * Lm1 + Cm2 * Lm2);
    if (diagram_mu < 0.) diagram_mu = 0.;
    return (diagram_e + diagram_mu) * 1E-01 / s.energy;
}

// SSR pair production, atomic DCS in m^2 / GeV (pumas.c:14624-14671): the
// 12-point rule over t = ln(rho) on [ln(1 - rmax), 0].
__device__ double pair_ssr(const ElementConsts& c, const Globals& G, double mass, double K, double q) {
#pragma clang fp contract(off)
    if ((c.Z <= 0) || (c.A <= 0) || (mass <= 0) || (K <= 0) || (q <= 0)) return 0.;
    if (q <= 4. * ELECTRON_MASS) return 0.;
    if (q >= K + c.p_qoff) return 0.;
    const double gamma = 1. + K / mass;
    const double x0 = 1. - 4. * ELECTRON_MASS / q;
    const double x1 = 1. - 6. / (gamma * (gamma - q / mass));
    if (!((x0 > 0) && (x1 > 0))) return 0.;
    const double rmax = sqrt(x0) * x1;
    const double ri = 1. - rmax;
    if ((ri <= 0.) || (ri >= 1.)) return 0.;
    const double tmin = log(ri);
    // the DDCS's own range check (pumas.c:14469-14471) holds at every node:
    // 4 m_e < q < K + p_qoff
    const PairQ s = pair_setup(c, mass, K, q);
    double I = 0.;
    for (int i = 0; i < 12; i++) {
        const double rho = exp(GQ12_X[i] * tmin);
        I -= pair_d2_ssr(s, G, rho) * rho * GQ12_W[i] * tmin;
    }
    return (I < 0) ? 0. : I;
}

// ALLM97 proton structure function F2 (pumas.c:14690-14745); lq0 is
// log(Q0^2 / Lambda^2).
__device__ double f2p_allm97(double x, double Q2, double lq0) {
#pragma clang fp contract(off)
    const double m02 = 0.31985, mP2 = 49.457, mR2 = 0.15052, Q02 = 0.52544, Lambda2 = 0.06527;
    const double M2 = NUCLEON_MASS * NUCLEON_MASS;
    const double W2 = M2 + Q2 * (1.0 / x - 1.0);
    const double t = log(log((Q2 + Q02) / Lambda2) / lq0);
    const double xP = (Q2 + mP2) / (Q2 + mP2 + W2 - M2);
    const double xR = (Q2 + mR2) / (Q2 + mR2 + W2 - M2);
    const double cP = 0.28067 + (0.28067 - 0.22291) * (1.0 / (1.0 + pow(t, 2.1979)) - 1.0);
    const double aP = -0.0808 + (-0.0808 - -0.44812) * (1.0 / (1.0 + pow(t, 1.1709)) - 1.0);
    const double bP = 0.36292 + 1.8917 * pow(t, 1.8439);
    const double cR = 0.80107 + 0.97307 * pow(t, 3.4942);
    const double aR = 0.58400 + 0.37888 * pow(t, 2.6063);
    const double bR = 0.01147 + 3.7582 * pow(t, 0.49338);
    const double l1x = log(1 - x);
    const double F2P = cP * exp(aP * log(xP) + bP * l1x);
    const double F2R = cR * exp(aR * log(xR) + bR * l1x);
    return Q2 / (Q2 + m02) * (F2P + F2R);
}

// DRSS photonuclear DDCS d2S / (dq dQ2) (pumas.c:14752-14785, 14987-15010).
__device__ double photo_d2_drss(const ElementConsts& c, const Globals& G, double ml, double K, double q, double Q2) {
#pragma clang fp contract(off)
    const double Z = c.Z, A = c.A;
    const double cf = 4 * PI * ALPHA_EM * ALPHA_EM * HBAR_C * HBAR_C;
    const double E = K + ml;
    const double y = q / E;
    const double x = 0.5 * Q2 / (NUCLEON_MASS * q);
    const double F2p = f2p_allm97(x, Q2, G.lq0);
    double shadowing;
    if (Z == 1.)
        shadowing = 1.;
    else if (x < 0.0014)
        shadowing = c.shadow_low;
    else if (x < 0.04)
        shadowing = exp((0.069 * log10(x) + 0.097) * c.lnA);
    else
        shadowing = 1.;
    const double F2A = F2p * shadowing * (Z + (A - Z) * (1.0 + x * (-1.85 + x * (2.45 + x * (-2.35 + x)))));
    const double dds = (1 - y + 0.5 * (1 - 2 * ml * ml / Q2) * (y * y + Q2 / (E * E)) / (1 + 0.)) / (Q2 * Q2) -
                       0.25 / (E * E * Q2);
    return cf * F2A * dds / q;
}

// DRSS photonuclear, atomic DCS in m^2 / GeV (pumas.c:15066-15146, mode 0):
// the 9-point rule on ln Q^2 over [Q2min, Q2max].
__device__ double photo_drss(const ElementConsts& c, const Globals& G, double ml, double K, double q) {
#pragma clang fp contract(off)
    if ((c.Z <= 0) || (c.A <= 0) || (ml <= 0) || (K <= 0)) return 0.;
    const double M = NUCLEON_MASS;
    const double mpi = PION_MASS;
    if ((q <= mpi + 0.5 * mpi * mpi / M) || (q >= K + ml - 0.5 * (M + ml * ml / M))) return 0.;
    const double E = K + ml;
    const double ml2 = ml * ml;
    const double Q2min = ml2 * (q * q - 0.5 * ml2) / (E * (E - q));
    const double Q2max = 2.0 * M * (q - mpi) - mpi * mpi;
    if ((Q2max < Q2min) | (Q2min < 0)) return 0.;
    const double lmin = log(Q2min);
    const double dx = log(Q2max) - lmin;
    double ds = 0.;
    for (int i = 0; i < 9; i++) {
        const double Q2 = exp(lmin + dx * GQ9_X[i]);
        const double tmp = photo_d2_drss(c, G, ml, K, q, Q2);
        if (tmp <= 0.) continue;
        ds += tmp * (GQ9_W[i] * (dx * Q2));
    }
    return (ds < 0.) ? 0. : ds;
}

// The m^2/kg DCS of process 0 (bremsstrahlung), 1 (pair production) or 2
// (photonuclear) as dcs_get(process) returns it (pumas.c:11018-11112).
__device__ double dcs(int process, const ElementConsts& c, const Globals& G, double mass, double K, double q) {
#pragma clang fp contract(off)
    double d;
    if (process == 0) {
        d = brems_ssr(c, G, mass, K, q);
    } else if (process == 1) {
        d = pair_ssr(c, G, mass, K, q);
    } else {
        const double qmin = PION_MASS + 0.5 * PION_MASS * PION_MASS / NUCLEON_MASS;
        const double qmax = K + mass - 0.5 * (NUCLEON_MASS + mass * mass / NUCLEON_MASS);
        if ((q < qmin) || (q > qmax)) return 0.;
        d = photo_drss(c, G, mass, K, q);
    }
    return d * 1E+03 * AVOGADRO_NUMBER * (mass + K) / c.A;
}

__global__ __launch_bounds__(256) void pumas_dcs_kernel(int process, ElementConsts c, Globals G, double mass,
                                                        const double* __restrict__ K, const double* __restrict__ q,
                                                        int64_t n, double* __restrict__ out) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        out[i] = dcs(process, c, G, mass, K[i], q[i]);
}

// One 64-lane workgroup per row of the launch's element block: block row
// b = (p n_blk + j) n_K + k is table row r = (p n_el + e0 + j) n_K + k.
__global__ __launch_bounds__(64) void pumas_table_kernel(Grid g, ElementBlock blk, Globals G, int32_t e0,
                                                         int32_t n_blk, int32_t n_el, double mass,
                                                         const double* __restrict__ Ks, int32_t n_K,
                                                         float* __restrict__ table, double* __restrict__ table64,
                                                         float* __restrict__ x_out, float* __restrict__ env,
                                                         const double* __restrict__ csn) {
#pragma clang fp contract(off)
    __shared__ double ys[MAX_X];
    __shared__ double ms[MAX_X];
    __shared__ int pos[MAX_X];
    const int64_t b = blockIdx.x;
    const int p = int(b / (int64_t(n_blk) * n_K));
    const int j = int((b / n_K) % n_blk);
    const int k = int(b % n_K);
    const int64_t r = (int64_t(p) * n_el + e0 + j) * n_K + k;
    const int n = g.n;
    const ElementConsts& c = blk.e[j];
    const double K = Ks[k];
    const int lane = threadIdx.x;
    if (csn && !(csn[r] > 0.)) {  // no cross section: an empty row (pumas.c:11427-11434)
        for (int i = lane; i < n; i += 64) {
            table[r * 2 * n + 2 * i] = 0.f;
            table[r * 2 * n + 2 * i + 1] = 0.f;
            if (table64) {
                table64[r * 2 * n + 2 * i] = 0.;
                table64[r * 2 * n + 2 * i + 1] = 0.;
            }
            if (b == 0 && x_out) x_out[i] = float(g.x[i]);
        }
        if (lane == 0 && env) {  // the reference leaves the pair as it was; here 0
            env[2 * r] = 0.f;
            env[2 * r + 1] = 0.f;
        }
        return;
    }

    for (int i = lane; i < n; i += 64) {
        const double q = K * exp(g.x[i]);
        const double d = dcs(p, c, G, mass, K, q);
        pos[i] = d > 0;
        ys[i] = d > 0 ? log(d) : 0.;
        if (b == 0 && x_out) x_out[i] = float(g.x[i]);
    }
    __syncthreads();
    // the first run of positive values (every lane scans the same LDS)
    int imin = n, imax = n - 1;
    for (int i = 0; i < n; i++) {
        if (pos[i]) {
            if (i < imin) imin = i;
        } else if (imin < n) {
            imax = i - 1;
            break;
        }
    }
    // PCHIP slopes over [imin, imax] (math_pchip_initialise with der = 0)
    const int N = imax - imin + 1;
    for (int i = lane; i < n; i += 64) {
        double m = 0.;
        if (N >= 1 && i >= imin && i <= imax) {
            const double* x = g.x + imin;
            const double* y = ys + imin;
            int j = i - imin;
            int nn = N;
            if (N >= 3) {
                if (j == 0) {
                    m = (y[1] - y[0]) / (x[1] - x[0]);
                    goto done;
                }
                x++;
                y++;
                j--;
                nn--;
            }
            if (nn == 1) {
                m = 0.;
            } else if (nn == 2) {
                m = (y[1] - y[0]) / (x[1] - x[0]);
            } else if (j == 0 || j == nn - 1) {
                // three-point end derivative (math_diff3, pumas.c:10010-10020)
                const int a = j, b = (j == 0) ? 1 : nn - 2, c = (j == 0) ? 2 : nn - 3;
                const double h1 = x[b] - x[a];
                const double h2 = x[c] - x[a];
                const double delta = h1 * h2 * (h2 - h1);
                const double c1 = h2 * h2 / delta;
                const double c2 = -h1 * h1 / delta;
                const double c0 = -(c1 + c2);
                m = c0 * y[a] + c1 * y[b] + c2 * y[c];
            } else {
                const double h1 = x[j] - x[j - 1];
                const double h2 = x[j + 1] - x[j];
                const double S1 = (y[j] - y[j - 1]) / h1;
                const double S2 = (y[j + 1] - y[j]) / h2;
                const double tmp = S1 * S2;
                if (tmp > 0) {
                    const double a = (h1 + 2 * h2) / (3 * (h1 + h2));
                    m = tmp / ((1 - a) * S1 + a * S2);
                } else {
                    m = 0.;
                }
            }
        }
    done:
        ms[i] = m;
        const bool in = i >= imin && i <= imax;
        float* row = table + r * 2 * n;
        row[2 * i] = in ? float(ys[i]) : 0.f;
        row[2 * i + 1] = in ? float(m) : 0.f;
        if (table64) {
            double* row64 = table64 + r * 2 * n;
            row64[2 * i] = in ? ys[i] : 0.;
            row64[2 * i + 1] = in ? m : 0.;
        }
    }
    __syncthreads();
    if (lane == 0 && env) {
        // envelope exponent (pumas.c:11462-11507)
        double qmin, qmax;
        if (p == 2) {
            qmin = PION_MASS + 0.5 * PION_MASS * PION_MASS / NUCLEON_MASS;
            qmax = K + mass - 0.5 * (NUCLEON_MASS + mass * mass / NUCLEON_MASS);
            qmin += qmin;
        } else {
            qmax = K + c.p_qoff;
            qmin = (p == 0) ? 0.5 * g.cutoff * K : 4. * ELECTRON_MASS * 4;
        }
        const double qrev = X_REVERSE * K;
        qmax *= 0.5;
        if (qmax > qrev) qmax = qrev;
        const double xmin = log(qmin / K), xmax = log(qmax / K);
        int ia0 = imax, ia1 = imin;
        for (int i = imin; i <= imax && i < n; i++) {
            if ((g.x[i] > xmin) && (g.x[i] < xmax)) {
                if (ia0 > i) ia0 = i;
                if (ia1 < i) ia1 = i;
            }
        }
        float alpha = 0.f;
        if ((ia0 < ia1) && (ia0 < n - 1)) {
            const int ia = (3 * ia0 + ia1) / 4;
            double a = -ms[ia];
            if (fabs(a - 1.) <= 0.1) a = 1.;
            alpha = float(a);
        }
        env[2 * r] = alpha;
        env[2 * r + 1] = 0.f;
    }
}

// Rows below the first positive envelope exponent of a (process, element)
// take that exponent (pumas.c:11745-11761). One thread per (process, element).
__global__ __launch_bounds__(64) void pumas_alpha_fill_kernel(float* __restrict__ env, int32_t n_pe, int32_t n_K) {
    const int pe = blockIdx.x * blockDim.x + threadIdx.x;
    if (pe >= n_pe) return;
    float* a = env + 2 * int64_t(pe) * n_K;
    int n0 = 0;
    float alpha = 0.f;
    for (int row = 0; row < n_K; row++) {
        if (a[2 * row] > 0.f) {
            n0 = row;
            alpha = a[2 * row];
            break;
        }
    }
    for (int row = 0; row < n0; row++) a[2 * row] = alpha;
}

// The m^2/kg model DCS for a whole wave that evaluates one (K, q): every lane
// computes the same set-up, the quadrature nodes are spread over the lanes
// (12 for pair production, 9 for photonuclear), and every lane then sums the
// gathered node terms in the serial loop's order, so the value is that of
// dcs() bit for bit. Must be called with the wave's 64 lanes active and the
// same arguments in every lane.
__device__ double dcs_wave(int process, const ElementConsts& c, const Globals& G, double mass, double K, double q) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    double d;
    if (process == 0) {
        d = brems_ssr(c, G, mass, K, q);
    } else if (process == 1) {
        if ((c.Z <= 0) || (c.A <= 0) || (mass <= 0) || (K <= 0) || (q <= 0)) return 0.;
        if (q <= 4. * ELECTRON_MASS) return 0.;
        if (q >= K + c.p_qoff) return 0.;
        const double gamma = 1. + K / mass;
        const double x0 = 1. - 4. * ELECTRON_MASS / q;
        const double x1 = 1. - 6. / (gamma * (gamma - q / mass));
        if (!((x0 > 0) && (x1 > 0))) return 0.;
        const double rmax = sqrt(x0) * x1;
        const double ri = 1. - rmax;
        if ((ri <= 0.) || (ri >= 1.)) return 0.;
        const double tmin = log(ri);
        const PairQ s = pair_setup(c, mass, K, q);
        double term = 0.;
        if (lane < 12) {
            const double rho = exp(GQ12_X[lane] * tmin);
            term = pair_d2_ssr(s, G, rho) * rho * GQ12_W[lane] * tmin;
        }
        double I = 0.;
        for (int i = 0; i < 12; i++) I -= __shfl(term, i, 64);
        d = (I < 0) ? 0. : I;
    } else {
        const double M = NUCLEON_MASS;
        const double mpi = PION_MASS;
        const double ml = mass;
        if ((q < mpi + 0.5 * mpi * mpi / M) || (q > K + ml - 0.5 * (M + ml * ml / M))) return 0.;
        if ((c.Z <= 0) || (c.A <= 0) || (ml <= 0) || (K <= 0)) return 0.;
        if ((q <= mpi + 0.5 * mpi * mpi / M) || (q >= K + ml - 0.5 * (M + ml * ml / M))) return 0.;
        const double E = K + ml;
        const double ml2 = ml * ml;
        const double Q2min = ml2 * (q * q - 0.5 * ml2) / (E * (E - q));
        const double Q2max = 2.0 * M * (q - mpi) - mpi * mpi;
        if ((Q2max < Q2min) | (Q2min < 0)) return 0.;
        const double lmin = log(Q2min);
        const double dx = log(Q2max) - lmin;
        double term = 0.;
        int skip = 1;
        if (lane < 9) {
            const double Q2 = exp(lmin + dx * GQ9_X[lane]);
            const double tmp = photo_d2_drss(c, G, ml, K, q, Q2);
            skip = tmp <= 0.;
            term = tmp * (GQ9_W[lane] * (dx * Q2));
        }
        double ds = 0.;
        for (int i = 0; i < 9; i++) {
            const double t = __shfl(term, i, 64);
            if (!__shfl(skip, i, 64)) ds += t;
        }
        d = (ds < 0.) ? 0. : ds;
    }
    return d * 1E+03 * AVOGADRO_NUMBER * (mass + K) / c.A;
}

// ---- the envelope maximum (dcs_tabulate_envelope_row, pumas.c:11526-11679) ----
//
// One wave per row: the reference's serial search (bisections that find where the DCS
// becomes positive at the range ends, then two bracketing bisections,
// math_find_minimum with algo 0, pumas.c:12172-12232 and 12416-12459, and for
// pair production at K >= 1e7 a logarithmic scan next to the upper end) of
// the maximum of DCS x^alpha over ln x, where DCS is the table interpolated as
// dcs_evaluate does without a context (pumas.c:11307-11398) and the model
// where the table does not apply. The rows read are those of the finished
// table. The search's control flow is the same in every lane of the wave
// (same row, same values), and the model evaluations spread their quadrature
// nodes over the lanes (dcs_wave), so rows do not diverge against each other. The reference's search reads the nodes from physics->table_DCS_x,
// which compute_dcs_table writes only after element 0 (pumas.c:11775-11783),
// so element 0's search sees zero nodes and takes the model everywhere; the
// kernel does the same.

struct EnvRow {
    const ElementConsts* c;
    const Globals* G;
    int process;
    double mass;
    const float* xt;    // the nodes as the search sees them (LDS)
    int n;
    // dcs_evaluate's energy bracket: K is the row's own energy at every
    // evaluation, so the two rows it reads (LDS copies) and their weights
    // are fixed per row
    const float* ra;  // row i0 (or the last row)
    const float* rb;  // row i0 + 1 (or the last row again)
    float h0, h1;
    bool below;  // K below the grid (never the case for a row's own energy)
};

__device__ __forceinline__ void bracketf(const float* table, float value, int* p1, int* p2) {
    while (true) {
        const int i3 = (*p1 + *p2) / 2;
        if (value >= table[i3])
            *p1 = i3;
        else
            *p2 = i3;
        if (*p2 - *p1 < 2) break;
    }
}

// table_index without a context (pumas.c:4249-4290, 4299-4306).
__device__ __forceinline__ int table_index(const double* table, int n_K, double value) {
    const int imax = n_K - 1;
    if (value < table[0]) return -1;
    if (value >= table[imax]) return imax;
    int i1 = 0, i2 = imax;
    while (true) {
        const int i3 = (i1 + i2) / 2;
        if (value >= table[i3])
            i1 = i3;
        else
            i2 = i3;
        if (i2 - i1 < 2) break;
    }
    return i1;
}

__device__ double dcs_interpolated(const EnvRow& a, double K, double q) {
#pragma clang fp contract(off)
    const double wj = K / (K + a.mass);
    const ElementConsts& c = *a.c;
    double qlow = 0., qmin, qmax;
    if (a.process == 2) {
        qmin = PION_MASS + 0.5 * PION_MASS * PION_MASS / NUCLEON_MASS;
        qmax = K + a.mass - 0.5 * (NUCLEON_MASS + a.mass * a.mass / NUCLEON_MASS);
        if ((q < qmin) || (q > qmax)) return 0.;
        qlow = 2 * qmin;
    } else {
        qmin = 4. * ELECTRON_MASS;
        qmax = K + c.p_qoff;
        if (a.process == 1)
            qlow = 4 * qmin;
        else
            qmin = 0.;
        if ((q < qmin) || (q > qmax)) return 0.;
    }
    if (q < qlow) return dcs_wave(a.process, c, *a.G, a.mass, K, q) * wj;
    const int nx = a.n;
    const float xmin = a.xt[0], xmax = a.xt[nx - 1];
    const float x = logf((float)(q / K));
    if ((x < xmin) || (x >= xmax)) return dcs_wave(a.process, c, *a.G, a.mass, K, q) * wj;
    int ix = 0, tmpi = nx - 1;
    bracketf(a.xt, x, &ix, &tmpi);
    if ((ix < 0) || (ix >= nx - 1)) return dcs_wave(a.process, c, *a.G, a.mass, K, q) * wj;
    if (a.below) return dcs_wave(a.process, c, *a.G, a.mass, K, q) * wj;
    float p00 = a.ra[2 * ix], m00 = a.ra[2 * ix + 1], p01 = a.ra[2 * ix + 2], m01 = a.ra[2 * ix + 3];
    float p10 = a.rb[2 * ix], m10 = a.rb[2 * ix + 1], p11 = a.rb[2 * ix + 2], m11 = a.rb[2 * ix + 3];
    const float h0 = a.h0, h1 = a.h1;
    if (((p00 == 0.f) && (m00 == 0.f)) || ((p01 == 0.f) && (m01 == 0.f)) || ((p10 == 0.f) && (m10 == 0.f)) ||
        ((p11 == 0.f) && (m11 == 0.f)))
        return dcs_wave(a.process, c, *a.G, a.mass, K, q) * wj;
    const float x0 = a.xt[ix];
    const float x1 = a.xt[ix + 1];
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

struct EnvObjective {
    const EnvRow& a;
    double K, alpha;
    __device__ double operator()(double lnx) const {
#pragma clang fp contract(off)
        const double x = exp(lnx);
        const double q = x * K;
        return -dcs_interpolated(a, K, q) * pow(x, alpha);
    }
};

template <class F>
__device__ int find_minimum_bisection(const F& f, double xa, double xb, double xc, const double* fb_p, double tol,
                                      int max_iter, double* fopt) {
#pragma clang fp contract(off)
    double fb = fb_p ? *fb_p : f(xb);
    tol *= fabs(xc - xa);
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

template <class F>
__device__ int find_minimum(const F& f, double xa, double xc, const double* fa_p, const double* fc_p, double tol,
                            int max_iter, double* fopt) {
#pragma clang fp contract(off)
    double fa = fa_p ? *fa_p : f(xa);
    double fc = fc_p ? *fc_p : f(xc);
    const double atol = tol * fabs(xc - xa);
    double xb = 0, fb = 0;
    int i;
    for (i = 0; i < max_iter; i++) {
        xb = 0.5 * (xa + xc);
        fb = f(xb);
        if ((fb < fa) && (fb < fc)) break;
        if (fabs(xc - xa) <= atol) {
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
        } else if (fc < fb) {
            fb = fc;
        }
        *fopt = fb;
        return -1;
    }
    const int status = find_minimum_bisection(f, xa, xb, xc, &fb, tol, max_iter - i, fopt);
    return (status < 0) ? -2 : status + i;
}

__device__ float envelope_max(const EnvRow& a, double K, double cutoff, float alpha, double cs) {
#pragma clang fp contract(off)
    constexpr double RESOLUTION = 1E-10;
    if (cs <= 0.) return 0.f;
    double qmin, qmax;
    if (a.process == 2) {  // dcs_get_range (pumas.c:11047-11056)
        qmin = PION_MASS + 0.5 * PION_MASS * PION_MASS / NUCLEON_MASS;
        qmax = K + a.mass - 0.5 * (NUCLEON_MASS + a.mass * a.mass / NUCLEON_MASS);
    } else {
        qmin = (a.process == 0) ? 0. : 4. * ELECTRON_MASS;
        qmax = K + a.c->p_qoff;
    }
    double xmin = qmin / K;
    if (xmin < cutoff) xmin = cutoff;
    xmin = log(xmin);
    double xmax = log(qmax / K);
    if (xmin >= xmax) return 0.f;
    const EnvObjective f{a, K, (double)alpha};
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
        if (K >= 1E+01) {
            const double xm = exp(xmax);
            double delta = RESOLUTION, x = xmax, fx = f1;
            while (xm > delta) {
                x = log(xm - delta);
                fx = f(x);
                if (fx != 0.)
                    break;
                else
                    delta += delta;
            }
            f1 = fx;
            x0 = xmax = x;
            x1 = log(xm - delta / 2);
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
    if ((a.process == 1) && (K >= 1E+07)) {  // a sharp peak next to xmax: a logarithmic scan
        constexpr int N_SCAN = 7;
        double xs[N_SCAN], fs[N_SCAN], fmin = 0.;
        int imin = -1;
        for (int i = 0; i < N_SCAN; i++) {
            xs[i] = xmax + log(1. - pow(10., double(-(i + 1))));
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

// One 64-lane workgroup per (process, block element j, kinetic energy k) row
// of the launch's element block; env[2 r + 1] of table row r.
__global__ __launch_bounds__(64) void pumas_envelope_kernel(Grid g, ElementBlock blk, Globals G, int32_t e0,
                                                            int32_t n_blk, int32_t n_el, double mass,
                                                            const double* __restrict__ Ks, int32_t n_K,
                                                            const float* __restrict__ table,
                                                            const double* __restrict__ csn, float* __restrict__ env) {
    __shared__ float xt[MAX_X];
    __shared__ float ra[2 * MAX_X], rb[2 * MAX_X];
    __shared__ ElementConsts sc;  // the row's element and the globals, in LDS for the search's calls
    __shared__ Globals sg;
    const int64_t b = blockIdx.x;
    const int p = int(b / (int64_t(n_blk) * n_K));
    const int j = int((b / n_K) % n_blk);
    const int k = int(b % n_K);
    const int e = e0 + j;
    const int64_t pe = int64_t(p) * n_el + e;
    const int64_t r = pe * n_K + k;
    for (int i = threadIdx.x; i < g.n; i += 64) xt[i] = (e == 0) ? 0.f : (float)g.x[i];
    if (threadIdx.x == 0) {
        sc = blk.e[j];
        sg = G;
    }
    // dcs_evaluate's energy bracket of the row's own energy (pumas.c:11352-11381)
    const double K = Ks[k];
    const float* rows = table + pe * n_K * 2 * g.n;
    const int imax = n_K - 1;
    int i0 = 0, ia = imax, ib = imax;
    float h0 = 1.f, h1 = 0.f;
    if (!((K >= Ks[imax]) || ((i0 = table_index(Ks, n_K, K)) >= imax)) && i0 >= 0) {
        ia = i0;
        ib = i0 + 1;
        const float K0 = (float)Ks[i0];
        const float K1 = (float)Ks[i0 + 1];
        h1 = logf(((float)K) / K0) / logf(K1 / K0);
        h0 = 1.f - h1;
    }
    for (int i = threadIdx.x; i < 2 * g.n; i += 64) {
        ra[i] = rows[int64_t(ia) * 2 * g.n + i];
        rb[i] = rows[int64_t(ib) * 2 * g.n + i];
    }
    __syncthreads();
    EnvRow a;
    a.c = &sc;
    a.G = &sg;
    a.process = p;
    a.mass = mass;
    a.n = g.n;
    a.xt = xt;
    a.ra = ra;
    a.rb = rb;
    a.h0 = h0;
    a.h1 = h1;
    a.below = i0 < 0;
    const float v = envelope_max(a, K, g.cutoff, env[2 * r], csn ? csn[r] : 1.);
    if (threadIdx.x == 0) env[2 * r + 1] = v;
}

// The grid (pumas.c:1548-1554, 11712-11737), on the host with its libm.
inline int n_table(double cutoff) {
    int n = N_REVERSE;
    if (cutoff < X_REVERSE) {
        int m = int(std::log10(X_REVERSE / cutoff) * PER_DECADE);
        if (m < N_MIN) m = N_MIN;
        n += m;
    }
    return n;
}

inline Globals make_globals() {
#pragma clang fp contract(off)
    Globals G;
    G.e1 = std::exp(1.);
    G.e05 = std::exp(0.5);
    G.e_16 = std::exp(-1. / 6.);
    G.e13 = std::exp(1. / 3.);
    G.e_13 = std::exp(-1. / 3.);
    G.lq0 = std::log(0.52544 / 0.06527);
    return G;
}

// Radiation logarithms by Z (pumas.c:13773-13802); 0 = use 182.7.
inline double host_rad_log(double Z) {
    static const double Lz[92] = {
        202.4, 151.9, 159.9, 172.3, 177.9, 178.3, 176.6, 173.4, 170.0, 165.8, 165.8, 167.1, 169.1, 170.8, 172.2,
        173.4, 174.3, 174.8, 175.1, 175.6, 176.2, 176.8, 0.0,   0.0,   0.0,   175.8, 0.0,   0.0,   173.1, 0.0,
        0.0,   173.0, 0.0,   0.0,   173.5, 0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   175.9, 0.0,   0.0,   0.0,
        0.0,   0.0,   0.0,   0.0,   177.4, 0.0,   0.0,   178.6, 0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,
        0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   177.6, 0.0,
        0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   178.0, 0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,   0.0,
        0.0,   179.8};
    const int i = int(Z);
    if (i >= 1 && i <= 92 && Lz[i - 1] > 0) return Lz[i - 1];
    return 182.7;
}

// The element's launch invariants, each in the reference's expression and
// order (pumas.c:13985-14016, 14301, 14478-14481, 14760, 14632-14634).
inline ElementConsts make_element(double Z, double A, double mass, const Globals& G) {
#pragma clang fp contract(off)
    ElementConsts c;
    c.Z = Z;
    c.A = A;
    c.rl = host_rad_log(Z);
    c.z13 = std::pow(Z, 1. / 3.);
    c.Z13m = std::pow(Z, -1. / 3.);
    c.Z23 = std::pow(Z, 2. / 3.);
    c.dn = 1.54 * std::pow(A, 0.27);
    c.lnA = std::log(A);
    c.shadow_low = std::exp(-0.1 * c.lnA);
    c.b_qoff = mass * (1. - 0.75 * SSR_SQRTE * c.z13);
    c.p_qoff = mass * (1. - 0.75 * PAIR_SQRTE * c.z13);
    const double m = mass * 1E+03;
    const double mu_qc = m / (SSR_MMU * G.e1 / c.dn);
    const double rho = std::sqrt(1. + 4. * mu_qc * mu_qc);
    const double log_rho = std::log((rho + 1.) / (rho - 1.));
    c.b_delta1 = std::log(mu_qc) + 0.5 * rho * log_rho;
    c.b_delta2 = std::log(mu_qc) + 0.25 * (3. * rho - rho * rho * rho) * log_rho + 2. * mu_qc * mu_qc;
    return c;
}

inline void make_grid(double cutoff, Grid& g) {
#pragma clang fp contract(off)
    const int n = n_table(cutoff);
    g.n = n;
    g.cutoff = cutoff;
    const int p = N_REVERSE;
    const int m = n - p;
    double xrev;
    if (m > 0) {
        xrev = X_REVERSE;
        const double x0 = std::log(cutoff);
        const double dx = std::log(xrev / cutoff) / (m - 1);
        for (int i = 0; i < m; i++) g.x[i] = x0 + i * dx;
    } else {
        xrev = cutoff;
    }
    const double dx = std::log((MAX_FRACTION - xrev) / DXMIN) / ((m > 0) ? p : p - 1);
    for (int i = 0; i < p; i++) g.x[n - 1 - i] = std::log(MAX_FRACTION - DXMIN * std::exp(i * dx));
}

}  // namespace pumas
}  // namespace noa_hip

using namespace noa_hip;

extern "C" {

int noa_pumas_n_table_dcs(double cutoff) {
    if (!(cutoff > 0.0 && cutoff < 1.0)) return fail(NOA_DCS_EINVAL, "cutoff must be in (0, 1) (got %g)", cutoff);
    const int n = pumas::n_table(cutoff);
    if (n > pumas::MAX_X) return fail(NOA_DCS_ERANGE, "cutoff %g needs %d grid nodes, more than %d", cutoff, n,
                                      pumas::MAX_X);
    return n;
}

int noa_pumas_dcs_f64(int32_t process, double Z, double A, double mass, const double* K, const double* q, int64_t n,
                      double* out, hipStream_t stream) {
    if (process < 0 || process > 2) return fail(NOA_DCS_EINVAL, "process must be 0, 1 or 2 (got %d)", process);
    if (!(Z >= 1.0) || !(A > 0.0) || !(mass > 0.0)) return fail(NOA_DCS_EINVAL, "Z >= 1, A > 0 and mass > 0 required");
    if (n < 0) return fail(NOA_DCS_EINVAL, "n must be >= 0");
    if (n == 0) return ok();
    if (!K || !q || !out) return fail(NOA_DCS_EINVAL, "K, q and out must be non-NULL");
    const pumas::Globals G = pumas::make_globals();
    const pumas::ElementConsts c = pumas::make_element(Z, A, mass, G);
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(pumas::pumas_dcs_kernel, dim3(unsigned(blocks < 65536 ? blocks : 65536)), dim3(256), 0, stream,
                       int(process), c, G, mass, K, q, n, out);
    if (int rc = check_hip(hipGetLastError(), "pumas_dcs_kernel")) return rc;
    return ok();
}

int noa_pumas_dcs_table(const double* Z, const double* A, int32_t n_elements, double mass, double cutoff,
                        const double* K, int32_t n_K, float* table, float* x, float* envelope, double* table_f64,
                        const double* csn, hipStream_t stream) {
    const int n = noa_pumas_n_table_dcs(cutoff);
    if (n < 0) return n;
    if (n_elements < 0 || n_K < 0) return fail(NOA_DCS_EINVAL, "n_elements and n_K must be >= 0");
    if (!(mass > 0.0)) return fail(NOA_DCS_EINVAL, "mass must be > 0 (got %g)", mass);
    if (int64_t(3) * n_elements * n_K == 0) return ok();
    if (int64_t(3) * pumas::MAX_EL * n_K > int64_t(0x7fffffff)) return fail(NOA_DCS_ERANGE, "n_K too large (%d)", n_K);
    if (!Z || !A || !K || !table) return fail(NOA_DCS_EINVAL, "Z, A, K and table must be non-NULL");
    for (int e = 0; e < n_elements; e++)
        if (!(Z[e] >= 1.0) || !(A[e] > 0.0))
            return fail(NOA_DCS_EINVAL, "element %d: Z >= 1 and A > 0 required (got %g, %g)", e, Z[e], A[e]);
    pumas::Grid g;
    pumas::make_grid(cutoff, g);
    const pumas::Globals G = pumas::make_globals();
    for (int e0 = 0; e0 < n_elements; e0 += pumas::MAX_EL) {
        const int n_blk = n_elements - e0 < pumas::MAX_EL ? n_elements - e0 : pumas::MAX_EL;
        pumas::ElementBlock blk{};
        for (int j = 0; j < n_blk; j++) blk.e[j] = pumas::make_element(Z[e0 + j], A[e0 + j], mass, G);
        const unsigned rows = unsigned(3 * n_blk * n_K);
        hipLaunchKernelGGL(pumas::pumas_table_kernel, dim3(rows), dim3(64), 0, stream, g, blk, G, e0, n_blk,
                           n_elements, mass, K, n_K, table, table_f64, e0 == 0 ? x : nullptr, envelope, csn);
        if (int rc = check_hip(hipGetLastError(), "pumas_table_kernel")) return rc;
    }
    if (envelope) {
        const int n_pe = 3 * n_elements;
        hipLaunchKernelGGL(pumas::pumas_alpha_fill_kernel, dim3(unsigned((n_pe + 63) / 64)), dim3(64), 0, stream,
                           envelope, n_pe, n_K);
        if (int rc = check_hip(hipGetLastError(), "pumas_alpha_fill_kernel")) return rc;
        // the envelope maxima, over the finished rows and exponents
        for (int e0 = 0; e0 < n_elements; e0 += pumas::MAX_EL) {
            const int n_blk = n_elements - e0 < pumas::MAX_EL ? n_elements - e0 : pumas::MAX_EL;
            pumas::ElementBlock blk{};
            for (int j = 0; j < n_blk; j++) blk.e[j] = pumas::make_element(Z[e0 + j], A[e0 + j], mass, G);
            const unsigned rows = unsigned(3 * n_blk * n_K);
            hipLaunchKernelGGL(pumas::pumas_envelope_kernel, dim3(rows), dim3(64), 0, stream, g, blk, G, e0, n_blk,
                               n_elements, mass, K, n_K, table, csn, envelope);
            if (int rc = check_hip(hipGetLastError(), "pumas_envelope_kernel")) return rc;
        }
    }
    return ok();
}

}  // extern "C"
