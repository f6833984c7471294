// One training step of the DCS surrogate network in three launches:
//
//   1. step_fwd_bwd_kernel, one workgroup of 4 waves per 32 rows (one MFMA
//      N-tile); wave m owns output features 32m..32m+31 (M-tile m) of every
//      layer and holds only that M-tile's weight fragments. Between layers the
//      waves exchange their bf16 outputs, the next layer's B operand, through
//      LDS (double-buffered, one barrier per layer). It gathers its rows of
//      the dataset, runs the forward pass (layer 1 in FP32 on the VALU, layers
//      2-5 on bf16 MFMA), the masked squared-error loss and its gradient, and
//      back-propagates through the transposed weights down to the first
//      layer's pre-activation gradient, keeping each layer's SiLU' in
//      registers. Each wave transposes its 32 x 32 FP32 tiles through LDS to
//      (a) write the activations and pre-activation gradients that the hidden
//      weight gradients need as bf16 in row-group-major order (8 rows of one
//      feature per 16 bytes: an MFMA operand fragment is one coalesced load),
//      and (b) sum the workgroup's rows of the FP32 gradients of the biases,
//      the first layer and the output layer into one partial per workgroup.
//   2. step_grad_kernel, per (row split, layer, output M-tile): the hidden
//      weight gradients dW_l = dZ_l^T A_(l-1) on bf16 MFMA with the rows as the
//      K dimension, as partial sums per split; one more workgroup per split
//      sums the split's workgroup partials of the FP32 gradients.
//   3. step_update_kernel, per parameter: the sum over the splits (a fixed
//      order), Adam, and the bf16 forward and transposed fragments of the
//      updated weight.
//
// Rounding points: bf16 MFMA operands (activations, weights, back-propagated
// gradients, and both operands of the hidden weight gradients), FP32
// accumulation; the first layer, all bias gradients and the output layer's
// weight gradient in FP32 from the FP32 values; SiLU and SiLU' in FP32 with
// the hardware exp and reciprocal, as the inference kernel evaluates SiLU.
// Every sum has a fixed order, so a step is deterministic. Rows beyond the
// batch are padded with zero mask: their gradients are exactly zero.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "capi_util.hh"
#include "mlp_common.hh"
#include "mlp_state.hh"

namespace noa_hip {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// The per-workgroup FP32 gradient partials, NSMALL floats each:
//   [0, 4H)            biases of layers 1-4 (layer-major)
//   [4H, 4H + H IN)    W1 [f][j]
//   then W5 [p][i] (OUT H), then b5 (OUT).
constexpr int SM_B = 0;
constexpr int SM_W1 = 4 * H;
constexpr int SM_W5 = SM_W1 + H * IN;
constexpr int SM_B5 = SM_W5 + OUT * H;
static_assert(SM_B5 + OUT == NSMALL, "small-gradient layout");

__host__ __device__ inline int64_t small_param(int k) {
    if (k < SM_W1) return k < H ? OFF_B1 + k : HIDDEN_B[k / H - 1] + k % H;
    if (k < SM_W5) return OFF_W1 + (k - SM_W1);
    if (k < SM_B5) return OFF_W5 + (k - SM_W5);
    return OFF_B5 + (k - SM_B5);
}

struct StepBufs {
    uint16_t* AT[3];   // activations of layers 1-3, bf16 row-group-major [rows/8][H][8]
    uint16_t* dZT[3];  // pre-activation gradients of layers 2-4, same layout
    float* small;      // [workgroups][NSMALL]
    float* loss;       // [workgroups]
};

__device__ __forceinline__ bf16x8 pack8(const float* f) {
    bf16x8 r;
#pragma unroll
    for (int e = 0; e < 8; e++) r[e] = __bf16(f[e]);
    return r;
}

__device__ __forceinline__ bf16x8 frag_at(const uint16_t* __restrict__ base, int64_t idx) {
    return *reinterpret_cast<const bf16x8*>(base + idx * 8);
}

// SiLU of z (a = z s, s = sigmoid z, with the hardware exp and reciprocal, as
// the inference kernel) and its derivative s (1 + z (1 - s)).
__device__ __forceinline__ void silu_and_slope(float z, float& a, float& ds) {
    const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
    a = z * s;
    ds = s * fmaf(z, 1.0f - s, 1.0f);
}

// Workgroup barrier that orders LDS only: outstanding global loads (the next
// layer's fragments) and stores (the gradient operands) stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

constexpr int TS = 36;  // FP32 transpose tile row: 32 rows + 4, 16-byte aligned

struct StepShared {
    bf16x8 xb[2][8][64];       // next layer's B operand, double-buffered: [buffer][K-block][lane]
    float tile[4][2][32][TS];  // per wave: [forward / backward][feature][row]
    float x[32][4];            // the rows' features
    float dy[32][4];           // the rows' output gradients
    float op[4][32][4];        // output-layer partials per wave: [wave][row][output]
};

// Writes the lane's 16 values of a 32 x 32 tile (its row c = lane % 32,
// features d_row(lane, v)) into t[feature][row].
__device__ __forceinline__ void tile_put(float (*t)[TS], int lane, const float* a) {
#pragma unroll
    for (int v = 0; v < 16; v++) t[d_row(lane, v)][lane & 31] = a[v];
}

// Reads rows 16h..16h+15 of feature c (h = lane / 32, c = lane % 32).
__device__ __forceinline__ void tile_get(const float (*t)[TS], int lane, float* r) {
    const float4* p = reinterpret_cast<const float4*>(&t[lane & 31][16 * (lane >> 5)]);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float4 v = p[j];
        r[4 * j] = v.x;
        r[4 * j + 1] = v.y;
        r[4 * j + 2] = v.z;
        r[4 * j + 3] = v.w;
    }
}

// The lane's 16 rows of feature f as two bf16 row groups (8 rows of one
// feature per 16 bytes) of a [rows/8][H][8] buffer; g0 = the first group.
__device__ __forceinline__ void put_rowgroups(uint16_t* __restrict__ dst, int64_t g0, int f, const float* r) {
    *reinterpret_cast<bf16x8*>(dst + (g0 * H + f) * 8) = pack8(r);
    *reinterpret_cast<bf16x8*>(dst + ((g0 + 1) * H + f) * 8) = pack8(r + 8);
}

// Sum of the two lane halves (fixed order: half 0 + half 1).
__device__ __forceinline__ float halves(float v) {
    const float o = __shfl_xor(v, 32, 64);
    return (threadIdx.x & 32) ? o + v : v + o;
}

__device__ __forceinline__ void load_frags(bf16x8* w, const uint16_t* __restrict__ base, int64_t first, int lane) {
#pragma unroll
    for (int b = 0; b < 8; b++) w[b] = frag_at(base, first + b * 64 + lane);
}

__global__ __launch_bounds__(256) void step_fwd_bwd_kernel(const float* __restrict__ X, const float* __restrict__ T,
                                                           const float* __restrict__ Mk,
                                                           RowOrder order, int64_t base, int rows,
                                                           const float* __restrict__ P,
                                                           const uint16_t* __restrict__ frag,
                                                           const uint16_t* __restrict__ fragT, StepBufs buf,
                                                           float inv_n) {
    __shared__ StepShared sh;
    const int lane = threadIdx.x & 63, m = threadIdx.x >> 6;
    const int h = lane >> 5, c = lane & 31;
    const int64_t r0 = int64_t(blockIdx.x) * 32;
    const int64_t r = r0 + c;           // this lane's row (its column of the N-tile)
    const int64_t g0 = r0 / 8 + 2 * h;  // its first row group in the transposed buffers
    float* __restrict__ small = buf.small + int64_t(blockIdx.x) * NSMALL;

    float x[IN], t[OUT], mk[OUT];
    {
        // unconditional loads (row 0 for the padding), selected afterwards
        const bool valid = r < rows;
        const int64_t g = valid ? row_of(order, base + r) : 0;
        float xv[IN], tv[OUT], mv[OUT];
#pragma unroll
        for (int j = 0; j < IN; j++) xv[j] = X[g * IN + j];
#pragma unroll
        for (int p = 0; p < OUT; p++) {
            tv[p] = T[g * OUT + p];
            mv[p] = Mk[g * OUT + p];
        }
#pragma unroll
        for (int j = 0; j < IN; j++) x[j] = valid ? xv[j] : 0.f;
#pragma unroll
        for (int p = 0; p < OUT; p++) {
            t[p] = valid ? tv[p] : 0.f;
            mk[p] = valid ? mv[p] : 0.f;
        }
    }

    bf16x8 w[8], wn[8];  // this wave's fragments of the current and the next layer
    load_frags(w, frag, (0 * 4 + m) * 8 * 64, lane);
    // the output layer's (K-blocks 2m, 2m+1) and its transpose's (M-tile m)
    const bf16x8 w5a = frag_at(frag, FRAG_HIDDEN / 8 + (2 * m) * 64 + lane);
    const bf16x8 w5b = frag_at(frag, FRAG_HIDDEN / 8 + (2 * m + 1) * 64 + lane);
    const bf16x8 w5t = frag_at(fragT, FRAGT_HIDDEN / 8 + m * 64 + lane);
    // the hidden layers' biases, loaded once: a load issued just before the
    // MFMAs would make them wait for every store still in flight
    float bias[3][16];
#pragma unroll
    for (int li = 0; li < 3; li++)
#pragma unroll
        for (int v = 0; v < 16; v++) bias[li][v] = P[HIDDEN_B[li] + 32 * m + d_row(lane, v)];

    // ---- forward ----
    float ds[4][16];  // SiLU' of the 4 hidden layers, this wave's M-tile
    float a[16];
#pragma unroll
    for (int v = 0; v < 16; v++) {  // layer 1, FP32 (the layered kernel's FMA order)
        const int f = 32 * m + d_row(lane, v);
        const float z = fmaf(P[OFF_W1 + f * IN + 2], x[2],
                             fmaf(P[OFF_W1 + f * IN + 1], x[1], fmaf(P[OFF_W1 + f * IN], x[0], P[OFF_B1 + f])));
        silu_and_slope(z, a[v], ds[0][v]);
    }
    sh.xb[0][2 * m][lane] = pack8(a);
    sh.xb[0][2 * m + 1][lane] = pack8(a + 8);
    tile_put(sh.tile[m][0], lane, a);
    if (m == 0 && h == 0)  // for the first layer's gradient
#pragma unroll
        for (int j = 0; j < IN; j++) sh.x[c][j] = x[j];
    lds_barrier();
#pragma unroll
    for (int li = 0; li < 3; li++) {  // layers 2-4: W_(li+2), input buffer li % 2
        {
            float rr[16];
            tile_get(sh.tile[m][0], lane, rr);  // the input activations, for dW_(li+2)
            put_rowgroups(buf.AT[li], g0, 32 * m + c, rr);
        }
        if (li < 2)
            load_frags(wn, frag, ((li + 1) * 4 + m) * 8 * 64, lane);
        else  // the first back-propagated layer
            load_frags(wn, fragT, (2 * 4 + m) * 8 * 64, lane);
        f32x16 acc;
#pragma unroll
        for (int v = 0; v < 16; v++) acc[v] = bias[li][v];
#pragma unroll
        for (int b = 0; b < 8; b++) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[b], sh.xb[li & 1][b][lane], acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 16; v++) silu_and_slope(acc[v], a[v], ds[li + 1][v]);
        sh.xb[(li + 1) & 1][2 * m][lane] = pack8(a);
        sh.xb[(li + 1) & 1][2 * m + 1][lane] = pack8(a + 8);
        tile_put(sh.tile[m][0], lane, a);
        lds_barrier();
        if (li < 2)
#pragma unroll
            for (int b = 0; b < 8; b++) w[b] = wn[b];
    }
    // output layer: wave m takes K-blocks 2m, 2m+1; rows 0-3 of the tile are
    // the 4 outputs (lanes < 32, v < 4)
    {
        f32x16 o;
#pragma unroll
        for (int v = 0; v < 16; v++) o[v] = 0.f;
        o = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w5a, sh.xb[1][2 * m][lane], o, 0, 0, 0);
        o = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w5b, sh.xb[1][2 * m + 1][lane], o, 0, 0, 0);
        if (h == 0)
#pragma unroll
            for (int p = 0; p < OUT; p++) sh.op[m][c][p] = o[p];
    }
    lds_barrier();

    // ---- loss: d = m (Y - T), dY = 2 d / (rows OUT), loss += d^2 / (rows OUT) ----
    float dy[OUT];
    float ls = 0.f;
#pragma unroll
    for (int p = 0; p < OUT; p++) {
        const float y = P[OFF_B5 + p] + ((sh.op[0][c][p] + sh.op[1][c][p]) + (sh.op[2][c][p] + sh.op[3][c][p]));
        const float d = (y - t[p]) * mk[p];
        dy[p] = 2.0f * d * inv_n;
        ls += (h == 0) ? d * d * inv_n : 0.f;
    }
    if (m == 0) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) ls += __shfl_xor(ls, off, 64);
        if (lane == 0) buf.loss[blockIdx.x] = ls;
        if (h == 0)
#pragma unroll
            for (int p = 0; p < OUT; p++) sh.dy[c][p] = dy[p];
    }

    // ---- back-propagation ----
    {  // dA4 = W5^T dY (dY in K-block 0: slots h = 0, e < 4), dZ4 = dA4 SiLU'(Z4)
        float dv[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (h == 0)
#pragma unroll
            for (int p = 0; p < OUT; p++) dv[p] = dy[p];
        f32x16 da;
#pragma unroll
        for (int v = 0; v < 16; v++) da[v] = 0.f;
        da = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w5t, pack8(dv), da, 0, 0, 0);
        float dz[16];
#pragma unroll
        for (int v = 0; v < 16; v++) dz[v] = da[v] * ds[3][v];
        sh.xb[0][2 * m][lane] = pack8(dz);
        sh.xb[0][2 * m + 1][lane] = pack8(dz + 8);
        tile_put(sh.tile[m][1], lane, dz);
    }
    lds_barrier();
    {  // the output layer's FP32 gradients: dW5 [p][32m + c], b5
        float hr[16];
        tile_get(sh.tile[m][0], lane, hr);  // A4 of feature 32m + c, rows 16h..16h+15
        float s[OUT] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 16; j++)
#pragma unroll
            for (int p = 0; p < OUT; p++) s[p] = fmaf(sh.dy[16 * h + j][p], hr[j], s[p]);
#pragma unroll
        for (int p = 0; p < OUT; p++) s[p] = halves(s[p]);
        if (h == 0)
#pragma unroll
            for (int p = 0; p < OUT; p++) small[SM_W5 + p * H + 32 * m + c] = s[p];
        if (m == 0 && lane < OUT) {
            float sb = 0.f;
            for (int j = 0; j < 32; j++) sb += sh.dy[j][lane];
            small[SM_B5 + lane] = sb;
        }
    }
#pragma unroll
    for (int li = 2; li >= 0; li--) {  // W_(li+2)^T: dZ_(li+2) -> dA_(li+1) -> dZ_(li+1)
        {  // dZ_(li+2): bf16 row groups for dW_(li+2), FP32 sum for b_(li+2)
            float rr[16];
            tile_get(sh.tile[m][1], lane, rr);
            put_rowgroups(buf.dZT[li], g0, 32 * m + c, rr);
            float sb = 0.f;
#pragma unroll
            for (int j = 0; j < 16; j++) sb += rr[j];
            sb = halves(sb);
            if (h == 0) small[SM_B + (li + 1) * H + 32 * m + c] = sb;
        }
#pragma unroll
        for (int b = 0; b < 8; b++) w[b] = wn[b];
        if (li > 0) load_frags(wn, fragT, ((li - 1) * 4 + m) * 8 * 64, lane);
        f32x16 acc;
#pragma unroll
        for (int v = 0; v < 16; v++) acc[v] = 0.f;
        const int in = (2 - li) & 1;  // dZ4 went to buffer 0
#pragma unroll
        for (int b = 0; b < 8; b++) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[b], sh.xb[in][b][lane], acc, 0, 0, 0);
        float dz[16];
#pragma unroll
        for (int v = 0; v < 16; v++) dz[v] = acc[v] * ds[li][v];
        if (li > 0) {
            sh.xb[in ^ 1][2 * m][lane] = pack8(dz);
            sh.xb[in ^ 1][2 * m + 1][lane] = pack8(dz + 8);
        }
        tile_put(sh.tile[m][1], lane, dz);
        lds_barrier();
    }
    {  // the first layer, FP32: dW1 [32m + c][j] = sum dZ1 x_j, b1
        float rr[16];
        tile_get(sh.tile[m][1], lane, rr);
        float s[IN + 1] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 16; j++) {
#pragma unroll
            for (int k = 0; k < IN; k++) s[k] = fmaf(rr[j], sh.x[16 * h + j][k], s[k]);
            s[IN] += rr[j];
        }
#pragma unroll
        for (int k = 0; k <= IN; k++) s[k] = halves(s[k]);
        if (h == 0) {
#pragma unroll
            for (int k = 0; k < IN; k++) small[SM_W1 + (32 * m + c) * IN + k] = s[k];
            small[SM_B + 32 * m + c] = s[IN];
        }
    }
}

// Split `blockIdx.x` of the K (row) dimension, K-steps of 16 rows
// [s0, s1). blockIdx.y < 12: layer li = y / 4 (weights of layer li + 2) and
// output M-tile mo = y % 4; wave ni takes input N-tile ni; the 32 x 32 tile
// goes to ws[split][param]. Each wave issues the loads of 16 K-steps before
// their MFMAs. blockIdx.y >= 12: one thread per FP32 gradient (and one for
// the loss) sums the split's workgroup partials (2 K-steps per workgroup),
// the loss to ws[STEP_SPLITS NPARAM + split].
constexpr int SMALL_BLOCKS = (NSMALL + 1 + 255) / 256;
constexpr int KCHUNK = 16;

__global__ __launch_bounds__(256) void step_grad_kernel(StepBufs buf, int ksteps, int steps_per_split,
                                                        float* __restrict__ ws) {
    const int split = blockIdx.x;
    const int s0 = split * steps_per_split;
    const int s1 = min(ksteps, s0 + steps_per_split);
    float* __restrict__ out = ws + int64_t(split) * NPARAM;
    if (blockIdx.y >= 12) {
        const int k = (blockIdx.y - 12) * blockDim.x + threadIdx.x;
        if (k > NSMALL) return;
        const int w0 = s0 / 2, w1 = s1 / 2;
        const bool is_loss = k == NSMALL;
        const float* src = is_loss ? buf.loss : buf.small + k;
        const int64_t stride = is_loss ? 1 : NSMALL;
        float g = 0.f;
        for (int wg = w0; wg < w1; wg += 8) {  // 8 loads in flight, then the fixed-order sum
            float a[8];
#pragma unroll
            for (int u = 0; u < 8; u++) a[u] = wg + u < w1 ? src[int64_t(wg + u) * stride] : 0.f;
#pragma unroll
            for (int u = 0; u < 8; u++) g += a[u];
        }
        if (is_loss)
            ws[int64_t(STEP_SPLITS) * NPARAM + split] = g;
        else
            out[small_param(k)] = g;
        return;
    }
    const int li = blockIdx.y / 4, mo = blockIdx.y % 4;
    const int lane = threadIdx.x & 63, ni = threadIdx.x >> 6;
    const int h = lane >> 5, c = lane & 31;
    const uint16_t* __restrict__ G = buf.dZT[li] + int64_t(32 * mo + c) * 8;
    const uint16_t* __restrict__ A = buf.AT[li] + int64_t(32 * ni + c) * 8;
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; v++) acc[v] = 0.f;
    for (int k0 = s0; k0 < s1; k0 += KCHUNK) {
        bf16x8 ga[KCHUNK], ab[KCHUNK];
        const bf16x8 zero = {};
#pragma unroll
        for (int u = 0; u < KCHUNK; u++) {
            const bool in = k0 + u < s1;
            const int64_t g = (2 * int64_t(k0 + u) + h) * H * 8;
            ga[u] = in ? *reinterpret_cast<const bf16x8*>(G + g) : zero;
            ab[u] = in ? *reinterpret_cast<const bf16x8*>(A + g) : zero;
        }
#pragma unroll
        for (int u = 0; u < KCHUNK; u++) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ga[u], ab[u], acc, 0, 0, 0);
    }
    const int64_t ow = HIDDEN_W[li];
#pragma unroll
    for (int v = 0; v < 16; v++) out[ow + int64_t(32 * mo + d_row(lane, v)) * H + 32 * ni + c] = acc[v];
}

// Sum of the split partials (fixed order) into grads; with `update`, Adam and
// the bf16 fragments of the updated weight (forward and transposed).
__global__ __launch_bounds__(256) void step_update_kernel(const float* __restrict__ ws, int splits,
                                                          float* __restrict__ grads, float* __restrict__ P,
                                                          float* __restrict__ m1, float* __restrict__ m2, int update,
                                                          float lr, float bc1, float bc2,
                                                          uint16_t* __restrict__ frag, uint16_t* __restrict__ fragT,
                                                          double* __restrict__ loss, double* __restrict__ loss_acc) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float part[STEP_SPLITS];
#pragma unroll
        for (int sp = 0; sp < STEP_SPLITS; sp++) part[sp] = sp < splits ? ws[int64_t(STEP_SPLITS) * NPARAM + sp] : 0.f;
        double l = 0.;
#pragma unroll
        for (int sp = 0; sp < STEP_SPLITS; sp++) l += double(part[sp]);
        *loss = l;
        if (loss_acc) *loss_acc += l;
    }
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < NPARAM; i += int64_t(gridDim.x) * blockDim.x) {
        float part[STEP_SPLITS];  // all loads in flight, then the fixed-order sum
#pragma unroll
        for (int s = 0; s < STEP_SPLITS; s++) part[s] = s < splits ? ws[int64_t(s) * NPARAM + i] : 0.f;
        float g = 0.f;
#pragma unroll
        for (int s = 0; s < STEP_SPLITS; s++) g += part[s];
        grads[i] = g;
        if (!update) continue;
        const float mi = b1 * m1[i] + (1.f - b1) * g;
        const float vi = b2 * m2[i] + (1.f - b2) * g * g;
        m1[i] = mi;
        m2[i] = vi;
        const float p = P[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
        P[i] = p;
        const uint16_t pb = __builtin_bit_cast(uint16_t, __bf16(p));
        int li = -1;
        int64_t off = 0;
        for (int l = 0; l < 3; l++)
            if (i >= HIDDEN_W[l] && i < HIDDEN_W[l] + H * H) {
                li = l;
                off = i - HIDDEN_W[l];
            }
        if (li >= 0) {
            const int o = int(off / H), k = int(off % H);  // W[o][k]: output o, input k
            frag[(((li * 4 + o / 32) * 8 + slot_b(k)) * 64 + 32 * slot_h(k) + o % 32) * 8 + slot_e(k)] = pb;
            fragT[(((li * 4 + k / 32) * 8 + slot_b(o)) * 64 + 32 * slot_h(o) + k % 32) * 8 + slot_e(o)] = pb;
        } else if (i >= OFF_W5 && i < OFF_W5 + OUT * H) {
            const int o = int((i - OFF_W5) / H), k = int((i - OFF_W5) % H);
            frag[FRAG_HIDDEN + (slot_b(k) * 64 + 32 * slot_h(k) + o) * 8 + slot_e(k)] = pb;
            fragT[FRAGT_HIDDEN + ((k / 32) * 64 + k % 32) * 8 + o] = pb;
        }
    }
}

// Both fragment blocks from the FP32 parameters, zero padding included.
__global__ void pack_all_kernel(const float* __restrict__ P, uint16_t* __restrict__ frag,
                                uint16_t* __restrict__ fragT) {
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < FRAG_TOTAL + FRAGT_TOTAL;
         idx += gridDim.x * blockDim.x) {
        float v;
        if (idx < FRAG_HIDDEN) {
            const int e = idx % 8, lane = (idx / 8) % 64, b = (idx / 512) % 8, m = (idx / 4096) % 4, li = idx / 16384;
            v = P[HIDDEN_W[li] + (32 * m + lane % 32) * H + frag_k(b, lane / 32, e)];
            frag[idx] = __builtin_bit_cast(uint16_t, __bf16(v));
        } else if (idx < FRAG_TOTAL) {
            const int j = idx - FRAG_HIDDEN;
            const int e = j % 8, lane = (j / 8) % 64, b = j / 512;
            const int row = lane % 32;
            v = row < OUT ? P[OFF_W5 + row * H + frag_k(b, lane / 32, e)] : 0.f;
            frag[idx] = __builtin_bit_cast(uint16_t, __bf16(v));
        } else if (idx < FRAG_TOTAL + FRAGT_HIDDEN) {
            const int j = idx - FRAG_TOTAL;
            const int e = j % 8, lane = (j / 8) % 64, b = (j / 512) % 8, m = (j / 4096) % 4, li = j / 16384;
            // rows: input feature 32 m + lane % 32; K: output feature frag_k(b, h, e)
            v = P[HIDDEN_W[li] + frag_k(b, lane / 32, e) * H + 32 * m + lane % 32];
            fragT[j] = __builtin_bit_cast(uint16_t, __bf16(v));
        } else {
            const int j = idx - FRAG_TOTAL - FRAGT_HIDDEN;
            const int e = j % 8, lane = (j / 8) % 64, m = j / 512;
            // K-block 0 of the 4 outputs: slots h = 0, e < 4
            v = (lane < 32 && e < OUT) ? P[OFF_W5 + e * H + 32 * m + lane] : 0.f;
            fragT[FRAGT_HIDDEN + j] = __builtin_bit_cast(uint16_t, __bf16(v));
        }
    }
}

int ensure_packed(noa_dcs_mlp* net, hipStream_t s) {
    if (net->packed_valid) return NOA_DCS_OK;
    hipLaunchKernelGGL(pack_all_kernel, dim3((FRAG_TOTAL + FRAGT_TOTAL + 255) / 256), dim3(256), 0, s, net->params,
                       net->packed, net->packedT);
    if (int rc = check_hip(hipGetLastError(), "pack_all_kernel")) return rc;
    net->packed_valid = true;
    return NOA_DCS_OK;
}

int train_step(noa_dcs_mlp* net, const float* X, const float* T, const float* Mk, const RowOrder& order, int64_t base,
               int rows, bool update, float lr, float bc1, float bc2, double* loss_acc, hipStream_t s) {
    if (rows < 1 || rows > net->max_rows) return fail(NOA_DCS_ERANGE, "rows must be in [1, max_rows]");
    if (int rc = ensure_packed(net, s)) return rc;
    MlpWork& w = net->work;
    const int rp = pad_rows(rows);
    const StepBufs b{{w.AT[0], w.AT[1], w.AT[2]}, {w.dZT[0], w.dZT[1], w.dZT[2]}, w.small, w.loss_partial};
    hipLaunchKernelGGL(step_fwd_bwd_kernel, dim3(unsigned(rp / 32)), dim3(256), 0, s, X, T, Mk, order, base, rows,
                       net->params, net->packed, net->packedT, b, 1.0f / float(int64_t(rows) * OUT));
    if (int rc = check_hip(hipGetLastError(), "step_fwd_bwd_kernel")) return rc;
    // splits hold whole workgroups of step_fwd_bwd_kernel (2 K-steps each)
    const int ksteps = rp / 16;
    int sps = (ksteps + STEP_SPLITS - 1) / STEP_SPLITS;
    sps += sps & 1;
    const int used = (ksteps + sps - 1) / sps;
    hipLaunchKernelGGL(step_grad_kernel, dim3(unsigned(used), 12 + SMALL_BLOCKS), dim3(256), 0, s, b, ksteps, sps,
                       w.ws);
    if (int rc = check_hip(hipGetLastError(), "step_grad_kernel")) return rc;
    hipLaunchKernelGGL(step_update_kernel, dim3(unsigned((NPARAM + 255) / 256)), dim3(256), 0, s, w.ws, used,
                       net->grads, net->params, net->adam_m, net->adam_v, int(update), lr, bc1, bc2, net->packed,
                       net->packedT, w.loss, loss_acc);
    if (int rc = check_hip(hipGetLastError(), "step_update_kernel")) return rc;
    return NOA_DCS_OK;
}

double step_loss(noa_dcs_mlp* net, hipStream_t s) {
    double l = 0.;
    (void)hipMemcpyAsync(&l, net->work.loss, sizeof(double), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    return l;
}

}  // namespace mlp
}  // namespace noa_hip
