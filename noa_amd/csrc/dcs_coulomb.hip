// Coulomb scattering of the PMS path on MI355X (src/noa/pms/dcs.hh:499-950):
//   coulomb_data       per K: centre-of-mass factors fCM [n][2], spin factor,
//                      screening parameters [n][9], inverse Wentzel path
//   coulomb_transport  per K: first two transport coefficients at a cutoff mu
//   hard_scattering    per K over a mixture of nel elements: cutoff angle mu0
//                      (Ridders' root, numerics.hh:154-214) and mean free
//                      path lb_h of elastic hard scatters
//   soft_scattering    per K: transverse transport of soft ionisation
//                      (closed form) plus photonuclear (a 102-node quadrature6
//                      of photonuclear * ratio * nu)
//
// The per-K functions are short and branchy, so they run one thread per K
// with the FP64 log/exp tables in LDS. Per-element values the reference
// recomputes in every call (pow(A, 1/3), pow(Z, 2/3), cos of the soft
// angle, the node values nu_j = 0.05 exp(t_j), which are the same for every
// K) are evaluated once on the host with the host libm. The photonuclear DCS
// at the soft-scattering nodes goes through the same compacting tile kernel
// as any (K, q) map; a per-K pass then sums the nodes in the reference's
// order. Every device function keeps the reference's operation order with
// FMA contraction off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "capi_util.hh"
#include "dcs_consts.hh"
#include "dcs_tile.hh"
#include "fastmath.hh"

namespace noa_hip {

constexpr double KIN_CUTOFF = 1E-9;    // physics.hh:78
constexpr double EHS_PATH_MAX = 1E+9;  // physics.hh:79
constexpr double EHS_OVER_MSC = 1E-4;  // physics.hh:80
constexpr int NSF = 9;                 // physics.hh:87

constexpr int SOFT_ITV = 17;               // ceil(100 / 6) subintervals (dcs.hh:936, numerics.hh:80)
constexpr int SOFT_NODES = 6 * SOFT_ITV;   // 102
constexpr int SOFT_CHUNK = 16384;          // K values per pass

// Host-evaluated per-element values (bit-identical to the reference's).
struct CoulombElement {
    double A, I, Zd, mass;
    int32_t Z;
    double Ma;       // A * 0.931494
    double M2;       // (mass + Ma)^2
    double two_Ma;   // 2 Ma
    double mmMa;     // mass (mass + Ma)
    double rM2;      // (mass / Ma)^2
    double R1sq;     // (1.02934 A^(1/3) + 0.435)^2
    double s0_c;     // 5.179587126e-12 Z^(2/3)
    double wp_c;     // A * 2.54910918e8
};

CoulombElement make_coulomb_element(double A, double I, int32_t Z, double mass) {
    CoulombElement c{};
    c.A = A;
    c.I = I;
    c.Z = Z;
    c.Zd = double(Z);
    c.mass = mass;
    c.Ma = A * 0.931494;
    c.M2 = mass + c.Ma;
    c.M2 *= c.M2;
    c.two_Ma = 2. * c.Ma;
    c.mmMa = mass * (mass + c.Ma);
    c.rM2 = mass / c.Ma;
    c.rM2 *= c.rM2;
    const double R1 = 1.02934 * std::pow(A, 1. / 3) + 0.435;
    c.R1sq = R1 * R1;
    c.s0_c = 5.179587126E-12 * std::pow(double(Z), 2. / 3.);
    c.wp_c = A * 2.54910918E+08;
    return c;
}

// ---- per-K closed forms (dcs.hh:499-602) ----

__device__ __forceinline__ void coulomb_data_one(double K, const CoulombElement& c, double* fCM, double* s,
                                                 double* fspin, double* invlambda, const fm::Tab& t) {
#pragma clang fp contract(off)
    const double mass = c.mass;
    // dcs.hh:499-521
    const double s_inv = 1. / sqrt(c.M2 + c.two_Ma * K);
    fCM[0] = (K + mass + c.Ma) * s_inv;
    double k0 = (K * c.Ma + c.mmMa) * s_inv - mass;
    if (k0 < KIN_CUTOFF) k0 = KIN_CUTOFF;
    const double etot = K + mass + c.Ma;
    const double beta2 = K * (K + 2. * mass) / (etot * etot);
    fCM[1] = sqrt(c.rM2 * (1. - beta2) + beta2);
    // dcs.hh:524-527 at k0
    const double e = k0 + mass;
    *fspin = k0 * (e + mass) / (e * e);
    // dcs.hh:540-602 at k0
    const double p2 = k0 * (k0 + 2. * mass);
    const double d = 5.8406E-02 / p2;
    s[1] = d / c.R1sq;
    s[2] = d / 4.;
    const double ZE = c.Zd * e;
    const double zeta2 = 5.3251346E-05 * (ZE * ZE) / p2;
    double cK;
    if (zeta2 > 1.) {
        const int32_t terms = 10 + c.Z;
        double f = 0.;
        for (int32_t i = 1; i <= terms; i++) f += zeta2 / (double(i) * (double(i * i) + zeta2));
        cK = fm::exp(f, t);
    } else {
        cK = fm::exp(1. - 1. / (1. + zeta2) + zeta2 * (0.2021 + zeta2 * (0.0083 * zeta2 - 0.0369)), t);
    }
    const double cM = 1. + 3.34 * zeta2;
    double r = k0 / e;
    r *= r;
    const double cc = r * cK + (1. - r) * cM;
    const double s0 = c.s0_c * cc / p2;
    s[0] = s0;
    const double d01 = 1. / (s0 - s[1]);
    const double d02 = 1. / (s0 - s[2]);
    const double d12 = 1. / (s[1] - s[2]);
    s[6] = d01 * d01 * d02 * d02;
    s[7] = d01 * d01 * d12 * d12;
    s[8] = d12 * d12 * d02 * d02;
    s[3] = 2. * s[6] * (d01 + d02);
    s[4] = 2. * s[7] * (d12 - d01);
    s[5] = -2. * s[8] * (d12 + d02);
    // dcs.hh:530-537 (Wentzel path at s0)
    const double dw = k0 * (k0 + 2. * mass) / (c.Zd * (k0 + mass));
    *invlambda = 1. / (c.wp_c * s0 * (1. + s0) * dw * dw);
}

// dcs.hh:624-674.
__device__ __forceinline__ void transport_one(const double* s, double fspin, double mu, double* G,
                                              const fm::Tab& t) {
#pragma clang fp contract(off)
    const double nuc = (s[1] < s[2]) ? s[1] : s[2];
    if (mu < 1E-08 * nuc) {
        const double L = fm::log(1. + mu / s[0], t);
        const double r = mu / (mu + s[0]);
        const double k = s[0] * (1. + s[0]);
        G[0] = k * (r / s[0] - fspin * (L - r));
        const double I2 = mu - s[0] * (r - 2. * L);
        G[1] = 2. * k * (L - r - fspin * I2);
        return;
    }
    const double half_mu2 = 0.5 * mu * mu;
    double g0 = 0., g1 = 0.;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double si = s[i];
        double r = mu / (mu + si);
        double L = fm::log(1. + mu / si, t);
        double mu1 = mu;
        const double I0 = r / si;
        const double J0 = L;
        const double I1 = L - r;
        r *= si;
        L *= si;
        const double J1 = mu1 - L;
        const double I2 = mu1 - 2. * L + r;
        L *= si;
        mu1 *= si;
        const double J2 = half_mu2 + L - mu1;
        g0 += s[3 + i] * (J0 - fspin * J1) + s[6 + i] * (I0 - fspin * I1);
        g1 += s[3 + i] * (J1 - fspin * J2) + s[6 + i] * (I1 - fspin * I2);
    }
    const double k = s[0] * (1. + s[0]) * s[1] * s[1] * s[2] * s[2];
    G[0] = g0 * k;
    G[1] = g1 * (2. * k);
}

// dcs.hh:697-735.
__device__ __forceinline__ double restricted_cs(double mu, double fspin, const double* s, const fm::Tab& t) {
#pragma clang fp contract(off)
    if (mu >= 1.) return 0.;
    const double nuc = (s[1] < s[2]) ? s[1] : s[2];
    if (mu < 1E-08 * nuc) {
        const double L = fm::log((s[0] + 1.) / (s[0] + mu), t);
        const double r = (1. - mu) / ((s[0] + mu) * (s[0] + 1.));
        const double k = s[0] * (1. + s[0]);
        return k * (r - fspin * (L - s[0] * r));
    }
    double cs = 0.;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double si = s[i];
        const double L = fm::log((si + 1.) / (si + mu), t);
        const double r = (1. - mu) / ((si + mu) * (si + 1.));
        const double I1 = L - si * r;
        const double J1 = mu - si * L;
        cs += s[3 + i] * (L - fspin * J1) + s[6 + i] * (r - fspin * I1);
    }
    const double k = s[0] * (1. + s[0]) * s[1] * s[1] * s[2] * s[2];
    return k * cs;
}

struct Mixture {
    const double* G;
    const double* fCM;
    const double* screen;
    const double* invlambda;
    const double* fspin;
    int64_t nel, nkin;
};

// dcs.hh:737-751 at kinetic-energy index i.
__device__ __forceinline__ double objective(double cs_h, double mu, const Mixture& m, int64_t i, const fm::Tab& t) {
#pragma clang fp contract(off)
    double cs = 0.;
    for (int64_t e = 0; e < m.nel; e++) {
        const int64_t o = e * m.nkin + i;
        cs += m.invlambda[o] * restricted_cs(mu, m.fspin[o], m.screen + NSF * o, t);
    }
    return cs - cs_h;
}

// numerics.hh:154-214, with the end values given; false where the
// reference returns std::nullopt.
__device__ bool ridders(double xa, double xb, double fa, double fb, double xtol, double rtol, uint32_t max_iter,
                        double cs_h, const Mixture& m, int64_t i, const fm::Tab& t, double* root) {
#pragma clang fp contract(off)
    if (fa * fb > 0) return false;
    if (fa == 0) { *root = xa; return true; }
    if (fb == 0) { *root = xb; return true; }
    const double tol = xtol + rtol * fmin(fabs(xa), fabs(xb));
    for (uint32_t it = 0; it < max_iter; it++) {
        double dm = 0.5 * (xb - xa);
        const double xm = xa + dm;
        const double fm_ = objective(cs_h, xm, m, i, t);
        double sgn = (fb > fa) ? 1. : -1.;
        double dn = sgn * dm * fm_ / sqrt(fm_ * fm_ - fa * fb);
        sgn = (dn > 0.) ? 1. : -1.;
        dn = fabs(dn);
        dm = fabs(dm) - 0.5 * tol;
        if (dn < dm) dm = dn;
        const double xn = xm - sgn * dm;
        const double fn = objective(cs_h, xn, m, i, t);
        if (fn * fm_ < 0.0) {
            xa = xn; fa = fn; xb = xm; fb = fm_;
        } else if (fn * fa < 0.0) {
            xb = xn; fb = fn;
        } else {
            xa = xn; fa = fn;
        }
        if (fn == 0.0 || fabs(xb - xa) < tol) { *root = xn; return true; }
    }
    return false;
}

// dcs.hh:753-838.
__device__ void hard_one(const Mixture& m, int64_t i, double max_mu0, double* mu0_out, double* lb_h_out,
                         const fm::Tab& t) {
#pragma clang fp contract(off)
    double invlb_m = 0., invlb1_m = 0., s_m_l = 0., s_m_h = 0.;
    for (int64_t e = 0; e < m.nel; e++) {
        const int64_t o = e * m.nkin + i;
        const double invlb = m.invlambda[o];
        const double scr = m.screen[NSF * o];
        invlb_m += invlb * m.G[2 * o];
        s_m_h += scr * invlb;
        s_m_l += invlb / scr;
        const double d = 1. / (m.fCM[2 * o] * (1. + m.fCM[1 + 2 * o]));
        invlb1_m += invlb * m.G[1 + 2 * o] * d * d;
    }
    const double lb_m = 1. / invlb_m;
    double lb_h = fmin(EHS_OVER_MSC / invlb1_m, EHS_PATH_MAX);
    double mu0;
    if (lb_m < lb_h) {
        const double s_m = (lb_h > 2. * lb_m) ? s_m_h * lb_m : 1. / (s_m_l * lb_m);
        mu0 = s_m * (lb_h - lb_m) / (s_m * lb_h + lb_m);
        const double cs_h = 1. / lb_h;
        double mu_max = fmin(4. * mu0, 1.);
        double mu_min = 0.25 * mu0;
        double fmax = objective(cs_h, mu_max, m, i, t), fmin_;
        if (fmax > 0.) {
            // as in the reference, mu0 and lb_h keep their estimates here
        } else {
            fmin_ = objective(cs_h, mu_min, m, i, t);
            if (fmin_ < 0.) {
                mu_max = mu_min;
                fmax = fmin_;
                mu_min = 0.;
                fmin_ = objective(cs_h, mu_min, m, i, t);
            }
            if (mu_min < max_mu0) {
                mu_max = fmin(mu_max, max_mu0);
                double root;
                if (ridders(mu_min, mu_max, fmin_, fmax, 1E-6 * mu0, 1E-6, 100, cs_h, m, i, t, &root)) mu0 = root;
            }
            mu0 = fmin(mu0, max_mu0);
            lb_h = objective(cs_h, mu0, m, i, t) + cs_h;
            lb_h = (lb_h <= 1. / EHS_PATH_MAX) ? EHS_PATH_MAX : 1. / lb_h;
        }
    } else {
        lb_h = lb_m;
        mu0 = 0;
    }
    *mu0_out = mu0;
    *lb_h_out = lb_h;
}

// dcs.hh:874-897.
__device__ __forceinline__ double transverse_ionisation(double K, const CoulombElement& c, const fm::Tab& t) {
#pragma clang fp contract(off)
    const double mass = c.mass, me = ELECTRON_MASS;
    const double p2 = K * (K + 2. * mass);
    const double E = K + mass;
    const double Wmax = 2. * me * p2 / (mass * mass + me * (me + 2. * E));
    const double W0 = 2. * p2 / me;
    const double mu_max = Wmax / W0;
    double mu3 = K * X_FRACTION / W0;
    if (mu3 > mu_max) mu3 = mu_max;
    const double mu2 = 0.62 * c.I / W0;
    if (mu2 >= mu3) return 0.;
    const double a0 = 0.5 * W0 / p2;
    const double a1 = -1. / Wmax;
    const double a2 = E * E / (W0 * p2);
    const double cs0 = 1.535336E-05 / c.A;
    return 2. * cs0 * c.Zd * (0.5 * a0 * (mu3 * mu3 - mu2 * mu2) + a1 * (mu3 - mu2) + a2 * fm::log(mu3 / mu2, t));
}

// dcs.hh:910-932: the mu-integrated ratio at one node.
__device__ __forceinline__ double photo_ratio(double K, double q, double mass, const fm::Tab& t) {
#pragma clang fp contract(off)
    const double E = K + mass;
    const double m02 = 0.4;
    const double q2 = q * q;
    const double tmax = 1.876544 * q;
    const double tmin = q2 * mass * mass / (E * (E - q));
    const double b1 = 1. / (1. - q2 / m02);
    const double c1 = 1. / (1. - m02 / q2);
    double L1 = b1 * fm::log((q2 + tmax) / (q2 + tmin), t);
    double L2 = c1 * fm::log((m02 + tmax) / (m02 + tmin), t);
    const double I0 = fm::log(tmax / tmin, t) - L1 - L2;
    L1 *= q2;
    L2 *= m02;
    const double I1 = L1 + L2;
    L1 *= q2;
    L2 *= m02;
    const double I2 = (tmax - tmin) * (b1 * q2 + c1 * m02) - L1 - L2;
    return (I1 * tmax - I2) / ((I0 * tmax - I1) * K * (K + 2. * mass));
}

// ---- kernels ----

#define NOA_FM_PROLOGUE()                                                                     \
    __shared__ __attribute__((aligned(fm::TABLE_LDS_ALIGN))) double fm_block[fm::TABLE_DOUBLES]; \
    load_fm_tables(fm_block);                                                                 \
    __syncthreads();                                                                          \
    const fm::Tab tab = fm::tab_from_block(fm_block)

#define NOA_GRID_LOOP(i, n) \
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); i += int64_t(gridDim.x) * blockDim.x)

__global__ __launch_bounds__(256) void coulomb_data_kernel(const double* __restrict__ K, int64_t n, CoulombElement c,
                                                           double* __restrict__ fCM, double* __restrict__ screen,
                                                           double* __restrict__ fspin,
                                                           double* __restrict__ invlambda) {
    NOA_FM_PROLOGUE();
    NOA_GRID_LOOP(i, n) {
        double f[2], s[NSF], sp, il;
        coulomb_data_one(K[i], c, f, s, &sp, &il, tab);
        fCM[2 * i] = f[0];
        fCM[2 * i + 1] = f[1];
#pragma unroll
        for (int k = 0; k < NSF; k++) screen[NSF * i + k] = s[k];
        fspin[i] = sp;
        invlambda[i] = il;
    }
}

__global__ __launch_bounds__(256) void coulomb_transport_kernel(const double* __restrict__ screen,
                                                                const double* __restrict__ fspin,
                                                                const double* __restrict__ mu, int64_t n_mu,
                                                                int64_t n, double* __restrict__ G) {
    NOA_FM_PROLOGUE();
    NOA_GRID_LOOP(i, n) {
        double s[NSF], g[2];
#pragma unroll
        for (int k = 0; k < NSF; k++) s[k] = screen[NSF * i + k];
        transport_one(s, fspin[i], mu[n_mu == 1 ? 0 : i], g, tab);
        G[2 * i] = g[0];
        G[2 * i + 1] = g[1];
    }
}

__global__ __launch_bounds__(256) void hard_scattering_kernel(Mixture m, double max_mu0, double* __restrict__ mu0,
                                                              double* __restrict__ lb_h) {
    NOA_FM_PROLOGUE();
    NOA_GRID_LOOP(i, m.nkin) {
        double a, b;
        hard_one(m, i, max_mu0, &a, &b, tab);
        mu0[i] = a;
        lb_h[i] = b;
    }
}

struct SoftNodes {
    double nu[SOFT_NODES];  // X_FRACTION exp(t_j), host libm
    double h;               // subinterval width in t
};

// Node j of K index i at j * nK + i (node-major: the per-K pass reads it
// coalesced).
__global__ __launch_bounds__(256) void soft_nodes_kernel(const double* __restrict__ K, int64_t nK, SoftNodes sn,
                                                         double* __restrict__ Kn, double* __restrict__ qn) {
    NOA_GRID_LOOP(idx, nK * SOFT_NODES) {
        const int j = int(idx / nK);
        const int64_t i = idx - int64_t(j) * nK;
        const double k = K[i];
        Kn[idx] = k;
        qn[idx] = sn.nu[j] * k;
    }
}

// numerics.hh:91-108 weights.
static __constant__ double SOFT_W[6] = {0.08566225, 0.18038079, 0.23395697, 0.23395697, 0.18038079, 0.08566225};

__global__ __launch_bounds__(256) void soft_reduce_kernel(const double* __restrict__ K, int64_t nK, SoftNodes sn,
                                                          const double* __restrict__ photo, CoulombElement c,
                                                          double* __restrict__ out) {
    NOA_FM_PROLOGUE();
    NOA_GRID_LOOP(i, nK) {
#pragma clang fp contract(off)
        const double k = K[i];
        double acc = 0.;
        for (int j = 0; j < SOFT_NODES; j++) {
            const double nu = sn.nu[j];
            const double q = nu * k;
            const double ratio = photo_ratio(k, q, c.mass, tab);
            const double f = photo[int64_t(j) * nK + i] * ratio * nu;
            acc += f * sn.h * SOFT_W[j % 6];
        }
        out[i] = transverse_ionisation(k, c, tab) + 2. * acc;
    }
}

unsigned grid_for(int64_t n) { return unsigned(std::max<int64_t>(1, std::min<int64_t>(65536, (n + 255) / 256))); }

}  // namespace noa_hip

using namespace noa_hip;

extern "C" {

int noa_dcs_coulomb_data_f64(const double* K, int64_t n, const noa_dcs_element* element, double mass, double* fCM,
                             double* screening, double* fspin, double* invlambda, hipStream_t stream) {
    if (int rc = check_element(element, mass)) return rc;
    if (n < 0) return fail(NOA_DCS_EINVAL, "n must be >= 0");
    if (n == 0) return ok();
    if (!K || !fCM || !screening || !fspin || !invlambda)
        return fail(NOA_DCS_EINVAL, "K, fCM, screening, fspin and invlambda must all be non-NULL");
    const CoulombElement c = make_coulomb_element(element->A, element->I, element->Z, mass);
    hipLaunchKernelGGL(coulomb_data_kernel, dim3(grid_for(n)), dim3(256), 0, stream, K, n, c, fCM, screening, fspin,
                       invlambda);
    if (int rc = check_hip(hipGetLastError(), "coulomb_data_kernel")) return rc;
    return ok();
}

int noa_dcs_coulomb_transport_f64(const double* screening, const double* fspin, const double* mu, int64_t n_mu,
                                  int64_t n, double* G, hipStream_t stream) {
    if (n < 0) return fail(NOA_DCS_EINVAL, "n must be >= 0");
    if (!(n_mu == 1 || n_mu == n)) return fail(NOA_DCS_EINVAL, "mu must hold 1 or n values (got %lld, n = %lld)",
                                               (long long)n_mu, (long long)n);
    if (n == 0) return ok();
    if (!screening || !fspin || !mu || !G) return fail(NOA_DCS_EINVAL, "screening, fspin, mu and G must be non-NULL");
    hipLaunchKernelGGL(coulomb_transport_kernel, dim3(grid_for(n)), dim3(256), 0, stream, screening, fspin, mu, n_mu,
                       n, G);
    if (int rc = check_hip(hipGetLastError(), "coulomb_transport_kernel")) return rc;
    return ok();
}

int noa_dcs_hard_scattering_f64(int64_t nel, int64_t nkin, const double* G, const double* fCM,
                                const double* screening, const double* invlambda, const double* fspin, double* mu0,
                                double* lb_h, hipStream_t stream) {
    if (nel < 1) return fail(NOA_DCS_EINVAL, "nel must be >= 1");
    if (nkin < 0) return fail(NOA_DCS_EINVAL, "nkin must be >= 0");
    if (nkin == 0) return ok();
    if (!G || !fCM || !screening || !invlambda || !fspin || !mu0 || !lb_h)
        return fail(NOA_DCS_EINVAL, "every array argument must be non-NULL");
    // physics.hh:81-84, with the host libm as in the reference
    const double max_mu0 = 0.5 * (1. - std::cos(1E+00 * M_PI / 180.));
    const Mixture m{G, fCM, screening, invlambda, fspin, nel, nkin};
    hipLaunchKernelGGL(hard_scattering_kernel, dim3(grid_for(nkin)), dim3(256), 0, stream, m, max_mu0, mu0, lb_h);
    if (int rc = check_hip(hipGetLastError(), "hard_scattering_kernel")) return rc;
    return ok();
}

int64_t noa_dcs_soft_scattering_workspace_bytes(int64_t n) {
    if (n <= 0) return 0;
    return std::min<int64_t>(n, SOFT_CHUNK) * SOFT_NODES * int64_t(3 * sizeof(double));
}

int noa_dcs_soft_scattering_f64(const double* K, int64_t n, const noa_dcs_element* element, double mass,
                                double* result, void* workspace, int64_t workspace_bytes, hipStream_t stream) {
    if (int rc = check_element(element, mass)) return rc;
    if (n < 0) return fail(NOA_DCS_EINVAL, "n must be >= 0");
    if (n == 0) return ok();
    if (!K || !result) return fail(NOA_DCS_EINVAL, "K and result must be non-NULL");
    if (!workspace || workspace_bytes < noa_dcs_soft_scattering_workspace_bytes(n))
        return fail(NOA_DCS_EINVAL, "workspace must hold noa_dcs_soft_scattering_workspace_bytes(n) bytes");
    // numerics.hh:72-108 on [ln 1e-6, 0] with min_points = 100 (dcs.hh:905-936)
    SoftNodes sn{};
    const double lb = std::log(1E-06);
    sn.h = (0. - lb) / double(uint32_t(SOFT_ITV));
    const double x6[6] = {0.03376524, 0.16939531, 0.38069041, 0.61930959, 0.83060469, 0.96623476};
    for (int j = 0; j < SOFT_NODES; j++)
        sn.nu[j] = X_FRACTION * std::exp(lb + sn.h * (double(uint32_t(j / 6)) + x6[j % 6]));
    const CoulombElement c = make_coulomb_element(element->A, element->I, element->Z, mass);
    const int64_t cap = std::min<int64_t>(n, SOFT_CHUNK);
    double* Kn = static_cast<double*>(workspace);
    double* qn = Kn + cap * SOFT_NODES;
    double* ph = qn + cap * SOFT_NODES;
    for (int64_t k0 = 0; k0 < n; k0 += SOFT_CHUNK) {
        const int64_t nK = std::min<int64_t>(SOFT_CHUNK, n - k0);
        const int64_t nn = nK * SOFT_NODES;
        hipLaunchKernelGGL(soft_nodes_kernel, dim3(grid_for(nn)), dim3(256), 0, stream, K + k0, nK, sn, Kn, qn);
        if (int rc = check_hip(hipGetLastError(), "soft_nodes_kernel")) return rc;
        const noa_dcs_outputs_f64 po{nullptr, nullptr, ph, nullptr};
        if (int rc = noa_dcs_eval_f64(Kn, qn, nn, element, mass, &po, stream)) return rc;
        hipLaunchKernelGGL(soft_reduce_kernel, dim3(grid_for(nK)), dim3(256), 0, stream, K + k0, nK, sn, ph, c,
                           result + k0);
        if (int rc = check_hip(hipGetLastError(), "soft_reduce_kernel")) return rc;
    }
    return ok();
}

}  // extern "C"
