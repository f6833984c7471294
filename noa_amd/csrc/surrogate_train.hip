// Training of the DCS surrogate network (see mlp_common.hh): the network
// object, the features, the layered forward (a bf16 MFMA GEMM with FP32
// accumulation, the FP32 first layer, bias + SiLU; used to check the fused
// kernels), and the host loop of minibatch epochs over a device-resident
// dataset, whose steps are the three fused kernels of surrogate_step.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "capi_util.hh"
#include "mlp_common.hh"

namespace noa_hip {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int GT = 64;  // GEMM output tile (GT x GT per 256-thread workgroup)
constexpr int GK = 32;  // K step staged in LDS

// C_z[M x N] = op(A)[M x K_z] op(B)[K_z x N] over the K range of split z;
// op(A)[m][k] = TA ? A[k lda + m] : A[m lda + k], op(B)[k][n] = TB ? B[n ldb + k] : B[k ldb + n].
// Out-of-range elements are staged as zeros.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                        const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                        int ldc, int k_chunk) {
    __shared__ __bf16 As[GT][GK + 8];
    __shared__ __bf16 Bs[GT][GK + 8];
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int wm = wave / 2, wn = wave % 2;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int kb = blockIdx.z * k_chunk;
    const int ke = min(K, kb + k_chunk);
    float* Cz = C + int64_t(blockIdx.z) * M * ldc;
    f32x16 acc;
    for (int v = 0; v < 16; v++) acc[v] = 0.f;
    for (int k0 = kb; k0 < ke; k0 += GK) {
        for (int t = tid; t < GT * GK; t += 256) {
            int mm, kk;
            if (TA) { mm = t % GT; kk = t / GT; } else { mm = t / GK; kk = t % GK; }
            const int gm = m0 + mm, gk = k0 + kk;
            float v = 0.f;
            if (gm < M && gk < ke) v = TA ? A[int64_t(gk) * lda + gm] : A[int64_t(gm) * lda + gk];
            As[mm][kk] = __bf16(v);
        }
        for (int t = tid; t < GT * GK; t += 256) {
            int nn, kk;
            if (TB) { nn = t / GK; kk = t % GK; } else { nn = t % GT; kk = t / GT; }
            const int gn = n0 + nn, gk = k0 + kk;
            float v = 0.f;
            if (gn < N && gk < ke) v = TB ? B[int64_t(gn) * ldb + gk] : B[int64_t(gk) * ldb + gn];
            Bs[nn][kk] = __bf16(v);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GK; ks += 16) {
            bf16x8 a, b;
            const int r = wm * 32 + lane % 32, c = wn * 32 + lane % 32, k = ks + 8 * (lane / 32);
#pragma unroll
            for (int e = 0; e < 8; e++) {
                a[e] = As[r][k + e];
                b[e] = Bs[c][k + e];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int gn = n0 + wn * 32 + lane % 32;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int gm = m0 + wm * 32 + d_row(lane, v);
        if (gm < M && gn < N) Cz[int64_t(gm) * ldc + gn] = acc[v];
    }
}

// First layer in FP32: Z1[r][f] = W1[f] . X[r] + b1[f]; H1 = silu(Z1).
__global__ void layer1_fwd_kernel(const float* __restrict__ X, const float* __restrict__ P, float* __restrict__ Z1,
                                  float* __restrict__ H1, int rows) {
    const int64_t total = int64_t(rows) * H;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t r = i / H;
        const int f = int(i % H);
        const float* w = P + OFF_W1 + f * IN;
        const float z = fmaf(w[2], X[r * IN + 2], fmaf(w[1], X[r * IN + 1], fmaf(w[0], X[r * IN], P[OFF_B1 + f])));
        Z1[i] = z;
        H1[i] = silu(z);
    }
}

// Z += b (broadcast over rows); H = silu(Z) when act, else H = Z.
__global__ void bias_act_kernel(float* __restrict__ Z, const float* __restrict__ b, float* __restrict__ Hout,
                                int rows, int cols, int act) {
    const int64_t total = int64_t(rows) * cols;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
        const float z = Z[i] + b[i % cols];
        Z[i] = z;
        Hout[i] = act ? silu(z) : z;
    }
}

__global__ void features_kernel(const double* __restrict__ K, const double* __restrict__ q, int64_t n, int Z,
                                float* __restrict__ X) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        float x[IN];
        features(K[i], q[i], Z, x);
        for (int j = 0; j < IN; j++) X[i * IN + j] = x[j];
    }
}

}  // namespace mlp
}  // namespace noa_hip

using namespace noa_hip;
using namespace noa_hip::mlp;

// ---- the network object (shared with surrogate_infer.hip through this struct) ----
#include "mlp_state.hh"

namespace {

inline unsigned grid_for(int64_t n, int64_t per_block = 256, int64_t cap = 4096) {
    return unsigned(std::max<int64_t>(1, std::min<int64_t>(cap, (n + per_block - 1) / per_block)));
}

template <bool TA, bool TB>
hipError_t gemm(hipStream_t s, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C,
                int ldc) {
    hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB>), dim3((N + GT - 1) / GT, (M + GT - 1) / GT, 1), dim3(256), 0, s, M,
                       N, K, A, lda, B, ldb, C, ldc, K);
    return hipGetLastError();
}

// Forward over `rows` inputs X: activations kept in the workspace for backward.
hipError_t forward(noa_dcs_mlp* net, const float* X, int rows, hipStream_t s) {
    const float* P = net->params;
    MlpWork& w = net->work;
    hipLaunchKernelGGL(layer1_fwd_kernel, dim3(grid_for(int64_t(rows) * H)), dim3(256), 0, s, X, P, w.Z[0], w.A[0],
                       rows);
    for (int l = 0; l < 3; l++) {
        hipError_t e = gemm<false, true>(s, rows, H, H, w.A[l], H, P + HIDDEN_W[l], H, w.Z[l + 1], H);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(bias_act_kernel, dim3(grid_for(int64_t(rows) * H)), dim3(256), 0, s, w.Z[l + 1],
                           P + HIDDEN_B[l], w.A[l + 1], rows, H, 1);
    }
    hipError_t e = gemm<false, true>(s, rows, OUT, H, w.A[3], H, P + OFF_W5, H, w.Y, OUT);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bias_act_kernel, dim3(grid_for(int64_t(rows) * OUT)), dim3(256), 0, s, w.Y, P + OFF_B5, w.Y,
                       rows, OUT, 0);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int64_t noa_dcs_mlp_num_params(void) { return NPARAM; }

int noa_dcs_mlp_features(const double* K, const double* q, int64_t n, int32_t Z, float* X, hipStream_t s) {
    if (n < 0) return fail(NOA_DCS_EINVAL, "n must be >= 0");
    if (n == 0) return ok();
    if (!K || !q || !X) return fail(NOA_DCS_EINVAL, "NULL argument");
    hipLaunchKernelGGL(features_kernel, dim3(grid_for(n, 256, 65536)), dim3(256), 0, s, K, q, n, int(Z), X);
    if (int rc = check_hip(hipGetLastError(), "features_kernel")) return rc;
    return ok();
}

int noa_dcs_mlp_create(uint64_t seed, int32_t max_rows, noa_dcs_mlp** out) {
    if (!out) return fail(NOA_DCS_EINVAL, "out is NULL");
    *out = nullptr;
    if (max_rows < 1 || max_rows > (1 << 20)) return fail(NOA_DCS_ERANGE, "max_rows must be in [1, 2^20]");
    noa_dcs_mlp* net = new noa_dcs_mlp();
    int rc = mlp_alloc(net, max_rows);
    if (rc) {
        mlp_free(net);
        return rc;
    }
    // Kaiming-uniform weights (SiLU gain ~ ReLU's), zero biases.
    std::vector<float> p(NPARAM, 0.f);
    std::mt19937_64 gen(seed);
    auto init = [&](int64_t off, int fan_out, int fan_in) {
        const float bound = std::sqrt(6.0f / float(fan_in));
        std::uniform_real_distribution<float> u(-bound, bound);
        for (int64_t i = 0; i < int64_t(fan_out) * fan_in; i++) p[off + i] = u(gen);
    };
    init(OFF_W1, H, IN);
    for (int l = 0; l < 3; l++) init(HIDDEN_W[l], H, H);
    init(OFF_W5, OUT, H);
    for (int64_t i = 0; i < int64_t(OUT) * H; i++) p[OFF_W5 + i] *= 0.1f;
    rc = check_hip(hipMemcpy(net->params, p.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy(params)");
    if (!rc) rc = check_hip(hipMemset(net->adam_m, 0, NPARAM * sizeof(float)), "hipMemset(adam)");
    if (!rc) rc = check_hip(hipMemset(net->adam_v, 0, NPARAM * sizeof(float)), "hipMemset(adam)");
    if (rc) {
        mlp_free(net);
        return rc;
    }
    *out = net;
    return ok();
}

int noa_dcs_mlp_destroy(noa_dcs_mlp* net) {
    if (net) mlp_free(net);
    return ok();
}

int noa_dcs_mlp_get_params(const noa_dcs_mlp* net, float* host, int64_t n) {
    if (!net || !host) return fail(NOA_DCS_EINVAL, "NULL argument");
    if (n != NPARAM) return fail(NOA_DCS_EINVAL, "n must be %lld", (long long)NPARAM);
    if (int rc = check_hip(hipMemcpy(host, net->params, NPARAM * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy"))
        return rc;
    return ok();
}

int noa_dcs_mlp_set_params(noa_dcs_mlp* net, const float* host, int64_t n) {
    if (!net || !host) return fail(NOA_DCS_EINVAL, "NULL argument");
    if (n != NPARAM) return fail(NOA_DCS_EINVAL, "n must be %lld", (long long)NPARAM);
    net->packed_valid = false;
    if (int rc = check_hip(hipMemcpy(net->params, host, NPARAM * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy"))
        return rc;
    return ok();
}

// Layered forward (the training kernels): Y[rows x 4] for X[rows x 3].
int noa_dcs_mlp_forward(noa_dcs_mlp* net, const float* X, int64_t rows, float* Y, hipStream_t s) {
    if (!net || !X || !Y) return fail(NOA_DCS_EINVAL, "NULL argument");
    if (rows < 0) return fail(NOA_DCS_EINVAL, "rows must be >= 0");
    for (int64_t r0 = 0; r0 < rows; r0 += net->max_rows) {
        const int n = int(std::min<int64_t>(net->max_rows, rows - r0));
        if (int rc = check_hip(forward(net, X + r0 * IN, n, s), "mlp forward")) return rc;
        if (int rc = check_hip(hipMemcpyAsync(Y + r0 * OUT, net->work.Y, size_t(n) * OUT * sizeof(float),
                                              hipMemcpyDeviceToDevice, s), "hipMemcpyAsync"))
            return rc;
    }
    return ok();
}

// One forward + backward on rows <= max_rows (the fused step kernels, no
// update); loss and gradients to the host.
int noa_dcs_mlp_forward_backward(noa_dcs_mlp* net, const float* X, const float* T, const float* Mk, int32_t rows,
                                 double* loss, float* grads_host, hipStream_t s) {
    if (!net || !X || !T || !Mk) return fail(NOA_DCS_EINVAL, "NULL argument");
    if (rows < 1 || rows > net->max_rows) return fail(NOA_DCS_ERANGE, "rows must be in [1, max_rows]");
    if (int rc = train_step(net, X, T, Mk, RowOrder{}, 0, rows, false, 0.f, 1.f, 1.f, nullptr, s)) return rc;
    if (loss) *loss = step_loss(net, s);
    if (grads_host)
        if (int rc = check_hip(hipMemcpy(grads_host, net->grads, NPARAM * sizeof(float), hipMemcpyDeviceToHost),
                               "hipMemcpy(grads)"))
            return rc;
    return ok();
}

// Minibatch Adam over the device dataset (X [n x 3], T and M [n x 4]) for
// `epochs` epochs; cosine learning rate from lr_max to lr_min. Each epoch
// visits n / batch batches in its own keyed order (RowOrder); the mean loss
// of each epoch's steps to loss_per_epoch (host, may be NULL), summed on the
// device, so that the loop does not synchronise until it ends.
int noa_dcs_mlp_train(noa_dcs_mlp* net, const float* X, const float* T, const float* Mk, int64_t n, int32_t batch,
                      int32_t epochs, float lr_max, float lr_min, uint64_t seed, float* loss_per_epoch, hipStream_t s) {
    if (!net || !X || !T || !Mk) return fail(NOA_DCS_EINVAL, "NULL argument");
    if (n < batch || batch < 1 || batch > net->max_rows) return fail(NOA_DCS_ERANGE, "need 1 <= batch <= min(n, max_rows)");
    if (n > (int64_t(1) << 30)) return fail(NOA_DCS_ERANGE, "n too large");
    if (epochs < 0) return fail(NOA_DCS_EINVAL, "epochs must be >= 0");
    if (epochs == 0) return ok();
    const int64_t steps_per_epoch = n / batch;
    const int64_t total_steps = steps_per_epoch * epochs;
    const float b1 = 0.9f, b2 = 0.999f;
    double* d_loss = nullptr;
    if (int rc = check_hip(hipMalloc(&d_loss, size_t(epochs) * sizeof(double)), "hipMalloc(loss)")) return rc;
    int rc = check_hip(hipMemsetAsync(d_loss, 0, size_t(epochs) * sizeof(double), s), "hipMemsetAsync(loss)");
    for (int ep = 0; ep < epochs && !rc; ep++) {
        const RowOrder order = epoch_order(seed, ep, n);
        for (int64_t st = 0; st < steps_per_epoch && !rc; st++) {
            const int64_t step = int64_t(ep) * steps_per_epoch + st;
            const float cosf_ = 0.5f * (1.f + std::cos(float(M_PI) * float(step) / float(std::max<int64_t>(1, total_steps))));
            const float lr = lr_min + (lr_max - lr_min) * cosf_;
            net->adam_t++;
            const float bc1 = 1.f - std::pow(b1, float(net->adam_t));
            const float bc2 = 1.f - std::pow(b2, float(net->adam_t));
            rc = train_step(net, X, T, Mk, order, st * batch, batch, true, lr, bc1, bc2, d_loss + ep, s);
        }
    }
    std::vector<double> ep_loss(epochs, 0.);
    if (!rc)
        rc = check_hip(hipMemcpyAsync(ep_loss.data(), d_loss, size_t(epochs) * sizeof(double), hipMemcpyDeviceToHost, s),
                       "hipMemcpy(loss)");
    if (!rc) rc = check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    (void)hipFree(d_loss);
    if (!rc && loss_per_epoch)
        for (int ep = 0; ep < epochs; ep++) loss_per_epoch[ep] = float(ep_loss[ep] / double(steps_per_epoch));
    return rc ? rc : ok();
}

}  // extern "C"
