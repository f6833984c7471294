// Training of the DCS surrogate network (see mlp_common.hh) with hand-written
// kernels: a bf16 MFMA GEMM (FP32 in/out, FP32 accumulation, split-K for the
// weight gradients), the FP32 first layer, bias + SiLU forward/backward, the
// masked squared-error loss and Adam. The host loop runs minibatch epochs
// over a device-resident dataset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
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

// C[i] = sum_z P[z][i] for i < n.
__global__ void splitk_reduce_kernel(const float* __restrict__ P, float* __restrict__ C, int n, int splits) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int z = 0; z < splits; z++) s += P[int64_t(z) * n + i];
        C[i] = s;
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

// dW1[f][j] = sum_r dZ1[r][f] X[r][j], db1[f] = sum_r dZ1[r][f]; one block per f.
__global__ __launch_bounds__(256) void layer1_grad_kernel(const float* __restrict__ X, const float* __restrict__ dZ1,
                                                          float* __restrict__ G, int rows) {
    const int f = blockIdx.x;
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    for (int r = threadIdx.x; r < rows; r += blockDim.x) {
        const float d = dZ1[int64_t(r) * H + f];
        s[0] += d * X[int64_t(r) * IN];
        s[1] += d * X[int64_t(r) * IN + 1];
        s[2] += d * X[int64_t(r) * IN + 2];
        s[3] += d;
    }
    __shared__ float red[4][256];
    for (int j = 0; j < 4; j++) red[j][threadIdx.x] = s[j];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w)
            for (int j = 0; j < 4; j++) red[j][threadIdx.x] += red[j][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int j = 0; j < 3; j++) G[OFF_W1 + f * IN + j] = red[j][0];
        G[OFF_B1 + f] = red[3][0];
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

// dZ = dH * silu'(Z), silu'(z) = s (1 + z (1 - s)), s = sigmoid(z).
__global__ void silu_bwd_kernel(const float* __restrict__ dH, const float* __restrict__ Z, float* __restrict__ dZ,
                                int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const float z = Z[i];
        const float s = 1.0f / (1.0f + expf(-z));
        dZ[i] = dH[i] * s * (1.0f + z * (1.0f - s));
    }
}

// db[c] = sum_r dZ[r][c]; one block per column.
__global__ __launch_bounds__(256) void bias_grad_kernel(const float* __restrict__ dZ, float* __restrict__ db, int rows,
                                                        int cols) {
    const int c = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) s += dZ[int64_t(r) * cols + c];
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}

// Masked squared error, normalised by rows * OUT: dY = 2 m (Y - T) / (rows OUT);
// per-block partial sums of m (Y - T)^2 / (rows OUT) go to partial[blockIdx.x].
__global__ __launch_bounds__(256) void loss_kernel(const float* __restrict__ Y, const float* __restrict__ T,
                                                   const float* __restrict__ Mk, float* __restrict__ dY,
                                                   float* __restrict__ partial, int rows) {
    const int n = rows * OUT;
    const float inv = 1.0f / float(n);
    float s = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float d = (Y[i] - T[i]) * Mk[i];
        dY[i] = 2.0f * d * inv;
        s += d * d * inv;
    }
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
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

// Rows idx[i] of X (3 wide), T and M (4 wide) gathered into a contiguous batch.
__global__ void gather_kernel(const float* __restrict__ X, const float* __restrict__ T, const float* __restrict__ Mk,
                              const int32_t* __restrict__ idx, float* __restrict__ Xb, float* __restrict__ Tb,
                              float* __restrict__ Mb, int rows) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        const int64_t s = idx[r];
        for (int j = 0; j < IN; j++) Xb[r * IN + j] = X[s * IN + j];
        for (int j = 0; j < OUT; j++) {
            Tb[r * OUT + j] = T[s * OUT + j];
            Mb[r * OUT + j] = Mk[s * OUT + j];
        }
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
hipError_t gemm(hipStream_t s, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
                float* ws, int splits) {
    if (splits <= 1) {
        hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB>), dim3((N + GT - 1) / GT, (M + GT - 1) / GT, 1), dim3(256), 0, s,
                           M, N, K, A, lda, B, ldb, C, ldc, K);
        return hipGetLastError();
    }
    const int chunk = ((K + splits - 1) / splits + GK - 1) / GK * GK;
    const int sp = (K + chunk - 1) / chunk;
    hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB>), dim3((N + GT - 1) / GT, (M + GT - 1) / GT, sp), dim3(256), 0, s, M,
                       N, K, A, lda, B, ldb, ws, N, chunk);
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(grid_for(int64_t(M) * N)), dim3(256), 0, s, ws, C, M * N, sp);
    return hipGetLastError();
}

// Forward over `rows` inputs X: activations kept in the workspace for backward.
hipError_t forward(noa_dcs_mlp* net, const float* X, int rows, hipStream_t s) {
    const float* P = net->params;
    MlpWork& w = net->work;
    hipLaunchKernelGGL(layer1_fwd_kernel, dim3(grid_for(int64_t(rows) * H)), dim3(256), 0, s, X, P, w.Z[0], w.A[0],
                       rows);
    for (int l = 0; l < 3; l++) {
        hipError_t e = gemm<false, true>(s, rows, H, H, w.A[l], H, P + HIDDEN_W[l], H, w.Z[l + 1], H, nullptr, 1);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(bias_act_kernel, dim3(grid_for(int64_t(rows) * H)), dim3(256), 0, s, w.Z[l + 1],
                           P + HIDDEN_B[l], w.A[l + 1], rows, H, 1);
    }
    hipError_t e = gemm<false, true>(s, rows, OUT, H, w.A[3], H, P + OFF_W5, H, w.Y, OUT, nullptr, 1);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bias_act_kernel, dim3(grid_for(int64_t(rows) * OUT)), dim3(256), 0, s, w.Y, P + OFF_B5, w.Y,
                       rows, OUT, 0);
    return hipGetLastError();
}

// Backward from dY = dLoss/dY (in w.dY) into net->grads.
hipError_t backward(noa_dcs_mlp* net, const float* X, int rows, hipStream_t s) {
    const float* P = net->params;
    float* G = net->grads;
    MlpWork& w = net->work;
    const int splits = std::max(1, std::min(64, rows / 256));
    hipError_t e = gemm<true, false>(s, OUT, H, rows, w.dY, OUT, w.A[3], H, G + OFF_W5, H, w.splitk, splits);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bias_grad_kernel, dim3(OUT), dim3(256), 0, s, w.dY, G + OFF_B5, rows, OUT);
    // dA4 = dY W5
    e = gemm<false, false>(s, rows, H, OUT, w.dY, OUT, P + OFF_W5, H, w.dA, H, nullptr, 1);
    if (e != hipSuccess) return e;
    for (int l = 2; l >= 0; l--) {  // hidden layer l + 2: Z[l+1] = A[l] W^T + b
        hipLaunchKernelGGL(silu_bwd_kernel, dim3(grid_for(int64_t(rows) * H)), dim3(256), 0, s, w.dA, w.Z[l + 1], w.dZ,
                           int64_t(rows) * H);
        e = gemm<true, false>(s, H, H, rows, w.dZ, H, w.A[l], H, G + HIDDEN_W[l], H, w.splitk, splits);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(bias_grad_kernel, dim3(H), dim3(256), 0, s, w.dZ, G + HIDDEN_B[l], rows, H);
        e = gemm<false, false>(s, rows, H, H, w.dZ, H, P + HIDDEN_W[l], H, w.dA, H, nullptr, 1);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(silu_bwd_kernel, dim3(grid_for(int64_t(rows) * H)), dim3(256), 0, s, w.dA, w.Z[0], w.dZ,
                       int64_t(rows) * H);
    hipLaunchKernelGGL(layer1_grad_kernel, dim3(H), dim3(256), 0, s, X, w.dZ, G, rows);
    return hipGetLastError();
}

hipError_t loss_and_grad(noa_dcs_mlp* net, const float* T, const float* Mk, int rows, hipStream_t s) {
    MlpWork& w = net->work;
    hipLaunchKernelGGL(loss_kernel, dim3(LOSS_BLOCKS), dim3(256), 0, s, w.Y, T, Mk, w.dY, w.loss_partial, rows);
    return hipGetLastError();
}

double read_loss(noa_dcs_mlp* net, hipStream_t s) {
    float part[LOSS_BLOCKS];
    (void)hipMemcpyAsync(part, net->work.loss_partial, sizeof(part), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    double acc = 0;
    for (float p : part) acc += p;
    return acc;
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

// One forward + backward on rows <= max_rows; loss and gradients to the host.
int noa_dcs_mlp_forward_backward(noa_dcs_mlp* net, const float* X, const float* T, const float* Mk, int32_t rows,
                                 double* loss, float* grads_host, hipStream_t s) {
    if (!net || !X || !T || !Mk) return fail(NOA_DCS_EINVAL, "NULL argument");
    if (rows < 1 || rows > net->max_rows) return fail(NOA_DCS_ERANGE, "rows must be in [1, max_rows]");
    if (int rc = check_hip(forward(net, X, rows, s), "mlp forward")) return rc;
    if (int rc = check_hip(loss_and_grad(net, T, Mk, rows, s), "mlp loss")) return rc;
    if (int rc = check_hip(backward(net, X, rows, s), "mlp backward")) return rc;
    if (loss) *loss = read_loss(net, s);
    if (grads_host)
        if (int rc = check_hip(hipMemcpy(grads_host, net->grads, NPARAM * sizeof(float), hipMemcpyDeviceToHost),
                               "hipMemcpy(grads)"))
            return rc;
    return ok();
}

// Minibatch Adam over the device dataset (X [n x 3], T and M [n x 4]) for
// `epochs` epochs; cosine learning rate from lr_max to lr_min. Mean loss per
// epoch to loss_per_epoch (host, may be NULL).
int noa_dcs_mlp_train(noa_dcs_mlp* net, const float* X, const float* T, const float* Mk, int64_t n, int32_t batch,
                      int32_t epochs, float lr_max, float lr_min, uint64_t seed, float* loss_per_epoch, hipStream_t s) {
    if (!net || !X || !T || !Mk) return fail(NOA_DCS_EINVAL, "NULL argument");
    if (n < batch || batch < 1 || batch > net->max_rows) return fail(NOA_DCS_ERANGE, "need 1 <= batch <= min(n, max_rows)");
    if (n > (int64_t(1) << 30)) return fail(NOA_DCS_ERANGE, "n too large");
    if (epochs < 0) return fail(NOA_DCS_EINVAL, "epochs must be >= 0");
    MlpWork& w = net->work;
    std::vector<int32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 gen(seed);
    const int64_t steps_per_epoch = n / batch;
    const int64_t total_steps = steps_per_epoch * epochs;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int32_t* d_perm = nullptr;
    if (int rc = check_hip(hipMalloc(&d_perm, n * sizeof(int32_t)), "hipMalloc(perm)")) return rc;
    int rc = NOA_DCS_OK;
    for (int ep = 0; ep < epochs && !rc; ep++) {
        std::shuffle(perm.begin(), perm.end(), gen);
        rc = check_hip(hipMemcpyAsync(d_perm, perm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s), "hipMemcpy");
        double ep_loss = 0;
        int64_t sampled = 0;
        for (int64_t st = 0; st < steps_per_epoch && !rc; st++) {
            const int64_t step = int64_t(ep) * steps_per_epoch + st;
            hipLaunchKernelGGL(gather_kernel, dim3(grid_for(batch)), dim3(256), 0, s, X, T, Mk, d_perm + st * batch, w.Xb,
                               w.Tb, w.Mb, batch);
            rc = check_hip(forward(net, w.Xb, batch, s), "mlp forward");
            if (!rc) rc = check_hip(loss_and_grad(net, w.Tb, w.Mb, batch, s), "mlp loss");
            if (!rc) rc = check_hip(backward(net, w.Xb, batch, s), "mlp backward");
            const float cosf_ = 0.5f * (1.f + std::cos(float(M_PI) * float(step) / float(std::max<int64_t>(1, total_steps))));
            const float lr = lr_min + (lr_max - lr_min) * cosf_;
            net->adam_t++;
            const float bc1 = 1.f - std::pow(b1, float(net->adam_t));
            const float bc2 = 1.f - std::pow(b2, float(net->adam_t));
            hipLaunchKernelGGL(adam_kernel, dim3(grid_for(NPARAM)), dim3(256), 0, s, net->params, net->grads, net->adam_m,
                               net->adam_v, int(NPARAM), lr, b1, b2, eps, bc1, bc2);
            if (!rc) rc = check_hip(hipGetLastError(), "adam");
            if (loss_per_epoch && st % 16 == 0 && !rc) {  // sampled, each sample a synchronisation
                ep_loss += read_loss(net, s);
                sampled++;
            }
        }
        if (loss_per_epoch) loss_per_epoch[ep] = float(ep_loss / double(std::max<int64_t>(1, sampled)));
    }
    net->packed_valid = false;
    (void)hipStreamSynchronize(s);
    (void)hipFree(d_perm);
    return rc ? rc : ok();
}

}  // extern "C"
