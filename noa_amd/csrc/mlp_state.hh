// Device state of one surrogate network (parameters, Adam moments, training
// workspace, packed inference fragments). Shared by the training and the
// inference translation units; the C ABI sees it as an opaque handle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "capi_util.hh"
#include "mlp_common.hh"

namespace noa_hip {
namespace mlp {

constexpr int STEP_SPLITS = 32;  // row splits of the weight-gradient partials

// Rows are padded to a multiple of 64 in the training step.
inline int pad_rows(int rows) { return (rows + 63) / 64 * 64; }

// FP32 gradients summed per 32-row workgroup of the step: the biases of the 4
// hidden layers, W1, W5 and b5.
constexpr int NSMALL = 4 * H + H * IN + OUT * H + OUT;  // 1412

// The order in which an epoch visits the dataset: row p of the epoch is
// dataset row perm(p), a keyed bijection of [0, n) computed where it is used
// (no permutation array to build or copy). perm is a 4-round balanced Feistel
// network on 2 half_bits bits (2^(2 half_bits) >= n), round function the
// murmur3 32-bit finaliser of (half ^ round key), made a bijection of [0, n)
// by cycle walking: values >= n are permuted again, which ends within the
// cycle of the start value (at most 4 rounds of the network are expected,
// since the domain is below 4 n). shuffled == 0: the identity.
struct RowOrder {
    int64_t n = 0;
    uint32_t key[4] = {0, 0, 0, 0};
    int half_bits = 0;
    int shuffled = 0;
};

__host__ __device__ inline uint32_t fmix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

__host__ __device__ inline int64_t row_of(const RowOrder& o, int64_t p) {
    if (!o.shuffled) return p;
    const uint32_t mask = (1u << o.half_bits) - 1u;
    uint64_t y = uint64_t(p);
    do {
        uint32_t L = uint32_t(y >> o.half_bits), R = uint32_t(y) & mask;
        for (int i = 0; i < 4; i++) {
            const uint32_t t = L ^ (fmix32(R ^ o.key[i]) & mask);
            L = R;
            R = t;
        }
        y = (uint64_t(L) << o.half_bits) | R;
    } while (int64_t(y) >= o.n);
    return int64_t(y);
}

// The order of epoch `epoch` of a run seeded `seed` over n rows (n <= 2^30).
inline RowOrder epoch_order(uint64_t seed, int64_t epoch, int64_t n) {
    RowOrder o;
    o.n = n;
    o.shuffled = 1;
    int bits = 1;
    while ((int64_t(1) << bits) < n) bits++;
    o.half_bits = (bits + 1) / 2;
    for (int i = 0; i < 4; i++)
        o.key[i] = fmix32(uint32_t(seed) ^ fmix32(uint32_t(seed >> 32) ^ fmix32(uint32_t(epoch) * 4u + uint32_t(i) + 1u)));
    return o;
}

struct MlpWork {
    // the layered forward's buffers, [rows x H] row-major
    float* Z[4] = {};  // pre-activations of the 4 hidden layers
    float* A[4] = {};  // their SiLU
    float* Y = nullptr;  // [rows x OUT]
    // the training step's, rows padded to 64
    uint16_t* AT[3] = {};   // activations of layers 1-3, bf16 [rows/8][H][8]
    uint16_t* dZT[3] = {};  // pre-activation gradients of layers 2-4, same layout
    float* small = nullptr;         // [rows/32][NSMALL]
    float* loss_partial = nullptr;  // one per 32 rows
    float* ws = nullptr;  // STEP_SPLITS partial gradients [STEP_SPLITS x NPARAM], then STEP_SPLITS losses
    double* loss = nullptr;  // the last step's loss
};

}  // namespace mlp
}  // namespace noa_hip

struct noa_dcs_mlp {
    int32_t max_rows = 0;
    float* params = nullptr;
    float* grads = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    int64_t adam_t = 0;
    noa_hip::mlp::MlpWork work;
    uint16_t* packed = nullptr;   // bf16 MFMA fragments of layers 2-5 (forward)
    uint16_t* packedT = nullptr;  // and of their transposes (back-propagation)
    bool packed_valid = false;    // both match params
    void* scratch = nullptr;  // per-call element table and work items of the surrogate evaluation
    size_t scratch_bytes = 0;
};

namespace noa_hip {
namespace mlp {

inline int mlp_alloc(noa_dcs_mlp* net, int max_rows) {
    net->max_rows = max_rows;
    const int rp = pad_rows(max_rows);
    const size_t rh = size_t(rp) * H * sizeof(float);
    const size_t ro = size_t(rp) * OUT * sizeof(float);
    int rc = NOA_DCS_OK;
    auto m = [&](void* p, size_t bytes) {
        if (!rc) rc = check_hip(hipMalloc(reinterpret_cast<void**>(p), bytes), "hipMalloc(mlp)");
    };
    m(&net->params, NPARAM * sizeof(float));
    m(&net->grads, NPARAM * sizeof(float));
    m(&net->adam_m, NPARAM * sizeof(float));
    m(&net->adam_v, NPARAM * sizeof(float));
    for (int l = 0; l < 4; l++) {
        m(&net->work.Z[l], rh);
        m(&net->work.A[l], rh);
    }
    m(&net->work.Y, ro);
    for (int l = 0; l < 3; l++) {
        m(&net->work.AT[l], size_t(rp) * H * sizeof(uint16_t));
        m(&net->work.dZT[l], size_t(rp) * H * sizeof(uint16_t));
    }
    m(&net->work.small, size_t(rp / 32) * NSMALL * sizeof(float));
    m(&net->work.loss_partial, size_t(rp / 32) * sizeof(float));
    m(&net->work.ws, size_t(STEP_SPLITS) * (NPARAM + 1) * sizeof(float));
    m(&net->work.loss, sizeof(double));
    m(&net->packed, FRAG_TOTAL * sizeof(uint16_t));
    m(&net->packedT, FRAGT_TOTAL * sizeof(uint16_t));
    return rc;
}

inline void mlp_free(noa_dcs_mlp* net) {
    void* ptrs[] = {net->params,       net->grads,        net->adam_m, net->adam_v,  net->work.Y,
                    net->work.small,   net->work.loss_partial, net->work.ws, net->work.loss, net->packed,
                    net->packedT,      net->scratch};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (int l = 0; l < 4; l++) {
        if (net->work.Z[l]) (void)hipFree(net->work.Z[l]);
        if (net->work.A[l]) (void)hipFree(net->work.A[l]);
    }
    for (int l = 0; l < 3; l++) {
        if (net->work.AT[l]) (void)hipFree(net->work.AT[l]);
        if (net->work.dZT[l]) (void)hipFree(net->work.dZT[l]);
    }
    delete net;
}

// Re-packs both fragment blocks from the FP32 parameters when they are stale
// (surrogate_step.hip).
int ensure_packed(noa_dcs_mlp* net, hipStream_t s);

// One training step on `rows` rows (rows <= max_rows) of the dataset X / T /
// M: row r of the batch is dataset row row_of(order, base + r). Forward,
// masked squared error, back-propagation and the parameter gradient into
// net->grads, the loss into net->work.loss (and added to *loss_acc, a device
// double, when not NULL); with `update`, one Adam step (learning rate lr,
// bias corrections bc1, bc2) that also re-packs the fragments
// (surrogate_step.hip).
int train_step(noa_dcs_mlp* net, const float* X, const float* T, const float* Mk, const RowOrder& order, int64_t base,
               int rows, bool update, float lr, float bc1, float bc2, double* loss_acc, hipStream_t s);

// The last step's loss (synchronises the stream).
double step_loss(noa_dcs_mlp* net, hipStream_t s);

}  // namespace mlp
}  // namespace noa_hip
