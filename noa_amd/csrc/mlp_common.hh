// The DCS surrogate network: inputs (log10 K, log10 nu, Z) -> 4 hidden layers
// of width 128 with SiLU -> 4 outputs (standardised ln(A * DCS) per process).
//
// Parameters live in one flat FP32 vector (master copy, trained with Adam):
//   W1[128][3] b1[128] | W2[128][128] b2 | W3 b3 | W4 b4 | W5[4][128] b5[4]
// (W[out][in], row-major). Hidden and output layers run on bf16 MFMA with
// FP32 accumulation; the 3-input first layer stays FP32 on the VALU, since
// bf16 inputs would quantise log10 K to ~0.02 decades.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace noa_hip {
namespace mlp {

constexpr int IN = 3;
constexpr int H = 128;
constexpr int OUT = 4;

constexpr int64_t OFF_W1 = 0;
constexpr int64_t OFF_B1 = OFF_W1 + H * IN;
constexpr int64_t OFF_W2 = OFF_B1 + H;
constexpr int64_t OFF_B2 = OFF_W2 + H * H;
constexpr int64_t OFF_W3 = OFF_B2 + H;
constexpr int64_t OFF_B3 = OFF_W3 + H * H;
constexpr int64_t OFF_W4 = OFF_B3 + H;
constexpr int64_t OFF_B4 = OFF_W4 + H * H;
constexpr int64_t OFF_W5 = OFF_B4 + H;
constexpr int64_t OFF_B5 = OFF_W5 + OUT * H;
constexpr int64_t NPARAM = OFF_B5 + OUT;  // 50564

constexpr int64_t HIDDEN_W[3] = {OFF_W2, OFF_W3, OFF_W4};
constexpr int64_t HIDDEN_B[3] = {OFF_B2, OFF_B3, OFF_B4};

// Input features, FP32 (shared by training data preparation and inference).
constexpr float LOGK_MID = 1.5f, LOGK_HALF = 4.5f;   // log10 K in [-3, 6]
constexpr float LOGNU_MID = -3.0f, LOGNU_HALF = 3.0f;  // log10 nu in [-6, 0]

__host__ __device__ inline void features(double K, double q, int Z, float* x) {
    x[0] = (float(log10(K)) - LOGK_MID) / LOGK_HALF;
    x[1] = (float(log10(q / K)) - LOGNU_MID) / LOGNU_HALF;
    x[2] = float(Z) / 46.0f - 1.0f;
}

__host__ __device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }

// MFMA 32x32x16 chaining. For D = A(32x16) * B(16x32) + C, lane l holds
// row l%32 of A and column l%32 of B with 8 K-elements each (A and B index K
// identically), and D column l%32 at rows 8(v/4) + 4(l/32) + v%4, v < 16.
// Layer outputs stay in registers as the next layer's B operand: K-block
// b = 2m + s of the next layer, element e, is output row 32m + 16s + 8(e/4) +
// 4h + e%4 of M-tile m (h = l/32), i.e. D value v = 8s + e. Weights are
// packed with the same input-feature permutation.
__host__ __device__ constexpr int frag_k(int b, int h, int e) {
    return 32 * (b / 2) + 16 * (b % 2) + 8 * (e / 4) + 4 * h + (e % 4);
}

__host__ __device__ constexpr int d_row(int lane, int v) { return 8 * (v / 4) + 4 * (lane / 32) + (v % 4); }

// Packed fragment block (bf16 as uint16): 3 hidden layers x 4 M-tiles x 8
// K-blocks x 64 lanes x 8, then the output layer (4 rows padded to 32) x 8
// K-blocks x 64 lanes x 8.
constexpr int FRAG_HIDDEN = 3 * 4 * 8 * 64 * 8;
constexpr int FRAG_OUT = 8 * 64 * 8;
constexpr int FRAG_TOTAL = FRAG_HIDDEN + FRAG_OUT;

// The transposed weights for back-propagation, W^T as the A operand: rows are
// a layer's input features (natural order, 4 M-tiles of 32), K its output
// features in the frag_k order, so that the gradient of a layer's output, held
// as that layer's forward output, is the B operand as it stands. Hidden
// layers as FRAG_HIDDEN; the output layer's 4 outputs fill K-block 0
// (slots h = 0, e < 4), 4 M-tiles x 64 lanes x 8.
constexpr int FRAGT_HIDDEN = 3 * 4 * 8 * 64 * 8;
constexpr int FRAGT_OUT = 4 * 64 * 8;
constexpr int FRAGT_TOTAL = FRAGT_HIDDEN + FRAGT_OUT;

// Inverse of frag_k: the (K-block, lane half, element) slot of logical K index k.
__host__ __device__ constexpr int slot_b(int k) { return 2 * (k / 32) + (k % 32) / 16; }
__host__ __device__ constexpr int slot_h(int k) { return (k % 8) / 4; }
__host__ __device__ constexpr int slot_e(int k) { return 4 * ((k % 16) / 8) + k % 4; }
static_assert(frag_k(slot_b(77), slot_h(77), slot_e(77)) == 77, "slot inverse");
static_assert(frag_k(slot_b(126), slot_h(126), slot_e(126)) == 126, "slot inverse");

}  // namespace mlp
}  // namespace noa_hip
