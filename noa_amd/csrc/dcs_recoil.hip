// DEL / CEL recoil integrals of the four DCS processes over a vector of
// kinetic energies (src/noa/pms/dcs.hh:89-130, 444-497, 955-1001):
//   del_p(K) = (1/E) sum_j dcs_p(K, q_j) q_j   h w_j,   cel with q_j^2,
// i.e. (1/E) int dcs dq and (1/E) int dcs q dq over t = ln q, with
// q_j = exp(t_j) at the 6-node Gauss-Legendre points of ceil(min_points/6)
// subintervals of [ln(K xlow), ln K] (numerics.hh:72-108). Ionisation below
// K = (m - m_e)^2 / (2 m_e) takes the reference's analytic form.
//
// The DCS map is the cost (180 evaluations per K and process at the
// reference's min_points = 180), so the nodes of a chunk of K values are
// materialised as (K, q) pairs and pushed through the same compacting tile
// kernel as any (K, q) map, all four processes at once; a per-K pass then
// sums the nodes in the reference's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "capi_util.hh"
#include "dcs_consts.hh"
#include "dcs_f64.hh"
#include "dcs_tile.hh"
#include "fastmath.hh"

namespace noa_hip {

// numerics.hh:97-100 (quadrature6 literals on [0, 1]).
static __constant__ double GL6_X[6] = {0.03376524, 0.16939531, 0.38069041, 0.61930959, 0.83060469, 0.96623476};
static __constant__ double GL6_W[6] = {0.08566225, 0.18038079, 0.23395697, 0.23395697, 0.18038079, 0.08566225};

constexpr int RECOIL_CHUNK = 65536;  // K values per pass

// Scratch of the no-allocation entry point noa_dcs_recoil_integrals_f64_noalloc:
// file-scope device memory, one copy per GPU the library runs on, held until
// the process ends. Host code never uses this name as an address (on the host
// it is a stand-in for the symbol): it takes the device address once per
// device from hipGetSymbolAddress.
constexpr int64_t RECOIL_SCRATCH_BYTES = int64_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[RECOIL_SCRATCH_BYTES];

struct RecoilGrid {
    double xlow;
    int32_t n_itv;  // subintervals
    int32_t nodes;  // 6 n_itv
};

// Node j of K: t = lb + h ((j / 6) + x_{j % 6}), lb = ln(K xlow), h = (ln K - lb) / n_itv.
__device__ __forceinline__ double node_t(double K, const RecoilGrid& g, int j, const fm::Tab& t, double* h_out) {
#pragma clang fp contract(off)
    const double lb = fm::log(K * g.xlow, t);
    const double ub = fm::log(K, t);
    const double h = (ub - lb) / double(uint32_t(g.n_itv));
    *h_out = h;
    return lb + h * (double(uint32_t(j / 6)) + GL6_X[j % 6]);
}

__global__ __launch_bounds__(256) void recoil_nodes_kernel(const double* __restrict__ K, int64_t nK, RecoilGrid g,
                                                           double* __restrict__ Kn, double* __restrict__ qn) {
    __shared__ __attribute__((aligned(fm::TABLE_LDS_ALIGN))) double fm_block[fm::TABLE_DOUBLES];
    load_fm_tables(fm_block);
    __syncthreads();
    const fm::Tab tab = fm::tab_from_block(fm_block);
    const int64_t total = nK * g.nodes;
    // node-major: node j of K index i at j * nK + i, so that the per-K pass
    // reads each node plane coalesced
    for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += int64_t(gridDim.x) * blockDim.x) {
        const int j = int(idx / nK);
        const int64_t i = idx - int64_t(j) * nK;
        const double k = K[i];
        double h;
        const double tj = node_t(k, g, j, tab, &h);
        Kn[idx] = k;
        qn[idx] = fm::exp(tj, tab);
    }
}

// dcs.hh:469-496 with the terms of :446-466.
__device__ __forceinline__ double analytic_ionisation(double K, double xlow, const DcsConsts& c, bool cel,
                                                      const fm::Tab& t) {
#pragma clang fp contract(off)
    const double me = ELECTRON_MASS;
    const double P2 = K * (K + c.i_2m);
    const double E = K + c.mass;
    const double Wmax = 2. * me * P2 / (c.i_m2 + me * (me + 2. * E));
    if (Wmax < X_FRACTION * K) return 0.;
    double Wmin = c.i_Wmin;
    const double qlow = K * xlow;
    if (qlow >= Wmin) Wmin = qlow;
    if (Wmax <= Wmin) return 0.;
    const double a0 = 0.5 / P2, a1 = -1. / Wmax, a2 = E * E / P2;
    double term;
    if (cel)
        term = 0.5 * a0 * (Wmax * Wmax - Wmin * Wmin) + a1 * (Wmax - Wmin) + a2 * fm::log(Wmax / Wmin, t);
    else
        term = a0 * (Wmax - Wmin) + a1 * fm::log(Wmax / Wmin, t) + a2 * (1. / Wmin - 1. / Wmax);
    return 1.535336E-05 * c.Zd / c.A * term;
}

// One thread per K: sequential sums over the nodes (the reference's order).
__global__ __launch_bounds__(256) void recoil_reduce_kernel(const double* __restrict__ K, int64_t nK, RecoilGrid g,
                                                            const double* __restrict__ qn, Outs<double> dcs,
                                                            DcsConsts c, Outs<double> del, Outs<double> cel) {
    __shared__ __attribute__((aligned(fm::TABLE_LDS_ALIGN))) double fm_block[fm::TABLE_DOUBLES];
    load_fm_tables(fm_block);
    __syncthreads();
    const fm::Tab tab = fm::tab_from_block(fm_block);
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nK; i += int64_t(gridDim.x) * blockDim.x) {
#pragma clang fp contract(off)
        const double k = K[i];
        double h;
        (void)node_t(k, g, 0, tab, &h);
        const double E = k + c.mass;
        const double* planes[4] = {dcs.b, dcs.p, dcs.ph, dcs.i};
        double* dst_del[4] = {del.b, del.p, del.ph, del.i};
        double* dst_cel[4] = {cel.b, cel.p, cel.ph, cel.i};
        const double m1 = c.mass - ELECTRON_MASS;
        const bool ion_analytic = k <= 0.5 * m1 * m1 / ELECTRON_MASS;
        for (int p = 0; p < 4; p++) {
            if (!dst_del[p] && !dst_cel[p]) continue;
            if (p == 3 && ion_analytic) {
                if (dst_del[p]) dst_del[p][i] = analytic_ionisation(k, g.xlow, c, false, tab);
                if (dst_cel[p]) dst_cel[p][i] = analytic_ionisation(k, g.xlow, c, true, tab);
                continue;
            }
            const double* d = planes[p] + i;
            const double* qq = qn + i;
            double sd = 0., sc = 0.;
            for (int j = 0; j < g.nodes; j++) {
                const double w = GL6_W[j % 6];
                const double qj = qq[int64_t(j) * nK];
                const double dq = d[int64_t(j) * nK] * qj;
                sd += dq * h * w;
                sc += dq * qj * h * w;
            }
            if (dst_del[p]) dst_del[p][i] = sd / E;
            if (dst_cel[p]) dst_cel[p][i] = sc / E;
        }
    }
}

}  // namespace noa_hip

using namespace noa_hip;

namespace {

int64_t nodes_of(int32_t min_points) { return int64_t((uint32_t(min_points) + 5u) / 6u) * 6; }

// Bytes of scratch per kinetic energy: the node pairs (K, q) and the four
// DCS planes, 6 doubles per node.
int64_t bytes_per_K(int32_t min_points) { return nodes_of(min_points) * int64_t(6 * sizeof(double)); }

int check_recoil_args(const double* K, int64_t n, double xlow, const noa_dcs_element* element, double mass,
                      int32_t min_points) {
    if (int rc = check_element(element, mass)) return rc;
    if (n < 0) return fail(NOA_DCS_EINVAL, "n must be >= 0");
    if (min_points < 1 || min_points > 6000) return fail(NOA_DCS_EINVAL, "min_points must be in [1, 6000]");
    if (!(xlow > 0.0 && xlow < 1.0)) return fail(NOA_DCS_EINVAL, "xlow must be in (0, 1) (got %g)", xlow);
    if (n > 0 && !K) return fail(NOA_DCS_EINVAL, "K is NULL");
    return NOA_DCS_OK;
}

// The passes over chunks of at most chunk_cap kinetic energies, with the
// node pairs and DCS planes in `work` (chunk_cap * bytes_per_K bytes).
int recoil_run(const double* K, int64_t n, double xlow, const noa_dcs_element* element, double mass,
               int32_t min_points, const noa_dcs_outputs_f64& D, const noa_dcs_outputs_f64& C, double* work,
               int64_t chunk_cap, hipStream_t stream) {
    const bool want[4] = {D.brems || C.brems, D.pair || C.pair, D.photo || C.photo, D.ion || C.ion};
    const RecoilGrid g{xlow, int32_t((uint32_t(min_points) + 5u) / 6u), int32_t(nodes_of(min_points))};
    double* Kn = work;
    double* qn = Kn + chunk_cap * g.nodes;
    double* planes = qn + chunk_cap * g.nodes;
    const DcsConsts c = make_consts(element->A, element->I, element->Z, mass);
    for (int64_t k0 = 0; k0 < n; k0 += chunk_cap) {
        const int64_t nK = std::min<int64_t>(chunk_cap, n - k0);
        const int64_t nn = nK * g.nodes;
        const unsigned grid = unsigned(std::min<int64_t>(65536, (nn + 255) / 256));
        hipLaunchKernelGGL(recoil_nodes_kernel, dim3(grid), dim3(256), 0, stream, K + k0, nK, g, Kn, qn);
        if (int rc = check_hip(hipGetLastError(), "recoil_nodes_kernel")) return rc;
        noa_dcs_outputs_f64 node_out{want[0] ? planes : nullptr, want[1] ? planes + nn : nullptr,
                                     want[2] ? planes + 2 * nn : nullptr, want[3] ? planes + 3 * nn : nullptr};
        if (int rc = noa_dcs_eval_f64(Kn, qn, nn, element, mass, &node_out, stream)) return rc;
        const Outs<double> d_out{node_out.brems, node_out.pair, node_out.photo, node_out.ion};
        auto shift = [k0](double* p) { return p ? p + k0 : nullptr; };
        const Outs<double> od{shift(D.brems), shift(D.pair), shift(D.photo), shift(D.ion)};
        const Outs<double> oc{shift(C.brems), shift(C.pair), shift(C.photo), shift(C.ion)};
        hipLaunchKernelGGL(recoil_reduce_kernel, dim3(unsigned((nK + 255) / 256)), dim3(256), 0, stream, K + k0, nK, g,
                           qn, d_out, c, od, oc);
        if (int rc = check_hip(hipGetLastError(), "recoil_reduce_kernel")) return rc;
    }
    return ok();
}

// Device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array at its own address).
int scratch_address(void** out) {
    static std::mutex mu;
    static void* addr[64] = {};
    int dev = 0;
    if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice")) return rc;
    if (dev < 0 || dev >= 64) return fail(NOA_DCS_ERANGE, "device ordinal %d beyond 63", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        if (int rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
            return rc;
        addr[dev] = p;
    }
    *out = addr[dev];
    return NOA_DCS_OK;
}

}  // namespace

extern "C" {

int64_t noa_dcs_recoil_workspace_bytes(int64_t n, int32_t min_points) {
    if (n <= 0 || min_points < 1) return 0;
    return std::min<int64_t>(n, RECOIL_CHUNK) * nodes_of(min_points) * int64_t(6 * sizeof(double));
}

int noa_dcs_recoil_integrals_f64(const double* K, int64_t n, double xlow, const noa_dcs_element* element,
                                 double mass, int32_t min_points, const noa_dcs_outputs_f64* del,
                                 const noa_dcs_outputs_f64* cel, void* workspace, int64_t workspace_bytes,
                                 hipStream_t stream) {
    if (int rc = check_recoil_args(K, n, xlow, element, mass, min_points)) return rc;
    const noa_dcs_outputs_f64 none{nullptr, nullptr, nullptr, nullptr};
    const noa_dcs_outputs_f64& D = del ? *del : none;
    const noa_dcs_outputs_f64& C = cel ? *cel : none;
    if (n == 0 || !(D.brems || D.pair || D.photo || D.ion || C.brems || C.pair || C.photo || C.ion)) return ok();
    if (!workspace || workspace_bytes < noa_dcs_recoil_workspace_bytes(n, min_points))
        return fail(NOA_DCS_EINVAL, "workspace must hold noa_dcs_recoil_workspace_bytes(n, min_points) bytes");
    return recoil_run(K, n, xlow, element, mass, min_points, D, C, static_cast<double*>(workspace),
                      std::min<int64_t>(n, RECOIL_CHUNK), stream);
}

int64_t noa_dcs_recoil_noalloc_scratch_bytes(void) { return RECOIL_SCRATCH_BYTES; }

int noa_dcs_recoil_integrals_f64_noalloc(const double* K, int64_t n, double xlow, const noa_dcs_element* element,
                                         double mass, int32_t min_points, const noa_dcs_outputs_f64* del,
                                         const noa_dcs_outputs_f64* cel, hipStream_t stream) {
    if (int rc = check_recoil_args(K, n, xlow, element, mass, min_points)) return rc;
    const noa_dcs_outputs_f64 none{nullptr, nullptr, nullptr, nullptr};
    const noa_dcs_outputs_f64& D = del ? *del : none;
    const noa_dcs_outputs_f64& C = cel ? *cel : none;
    if (n == 0 || !(D.brems || D.pair || D.photo || D.ion || C.brems || C.pair || C.photo || C.ion)) return ok();
    const int64_t chunk_cap = std::min<int64_t>(n, RECOIL_SCRATCH_BYTES / bytes_per_K(min_points));
    if (chunk_cap < 1)
        return fail(NOA_DCS_ERANGE, "min_points %d needs %lld scratch bytes per kinetic energy, more than the %lld held",
                    min_points, (long long)bytes_per_K(min_points), (long long)RECOIL_SCRATCH_BYTES);
    void* work = nullptr;
    if (int rc = scratch_address(&work)) return rc;
    return recoil_run(K, n, xlow, element, mass, min_points, D, C, static_cast<double*>(work), chunk_cap, stream);
}

}  // extern "C"
