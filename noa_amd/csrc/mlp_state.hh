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

constexpr int LOSS_BLOCKS = 64;
constexpr int SPLITK_MAX = 64;

struct MlpWork {
    float* Z[4] = {};  // pre-activations of the 4 hidden layers [rows x H]
    float* A[4] = {};  // their SiLU [rows x H]
    float* Y = nullptr;
    float* dY = nullptr;
    float* dA = nullptr;
    float* dZ = nullptr;
    float* splitk = nullptr;
    float* loss_partial = nullptr;
    float* Xb = nullptr;
    float* Tb = nullptr;
    float* Mb = nullptr;
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
    uint16_t* packed = nullptr;  // bf16 MFMA fragments of layers 2-5
    bool packed_valid = false;
    void* scratch = nullptr;  // per-call element table and work items of the surrogate evaluation
    size_t scratch_bytes = 0;
};

namespace noa_hip {
namespace mlp {

inline int mlp_alloc(noa_dcs_mlp* net, int max_rows) {
    net->max_rows = max_rows;
    const size_t rh = size_t(max_rows) * H * sizeof(float);
    const size_t ro = size_t(max_rows) * OUT * sizeof(float);
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
    m(&net->work.dY, ro);
    m(&net->work.dA, rh);
    m(&net->work.dZ, rh);
    m(&net->work.splitk, size_t(SPLITK_MAX) * H * H * sizeof(float));
    m(&net->work.loss_partial, LOSS_BLOCKS * sizeof(float));
    m(&net->work.Xb, size_t(max_rows) * IN * sizeof(float));
    m(&net->work.Tb, ro);
    m(&net->work.Mb, ro);
    m(&net->packed, FRAG_TOTAL * sizeof(uint16_t));
    return rc;
}

inline void mlp_free(noa_dcs_mlp* net) {
    void* ptrs[] = {net->params, net->grads, net->adam_m, net->adam_v, net->work.Y, net->work.dY, net->work.dA,
                    net->work.dZ, net->work.splitk, net->work.loss_partial, net->work.Xb, net->work.Tb, net->work.Mb,
                    net->packed, net->scratch};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (int l = 0; l < 4; l++) {
        if (net->work.Z[l]) (void)hipFree(net->work.Z[l]);
        if (net->work.A[l]) (void)hipFree(net->work.A[l]);
    }
    delete net;
}

}  // namespace mlp
}  // namespace noa_hip
