/*
 * noa_dcs.h - C ABI of the MI355X (gfx950) PMS differential cross-section library.
 *
 * Evaluates the four muon DCS processes of grinisrit/noa elementwise over
 * flat arrays of kinetic energy K (GeV) and energy transfer q (GeV) on the
 * GPU. Results are dsigma/dnu per unit mass (m^2/kg), as in the reference.
 *
 * Conventions (all entry points):
 *   - return 0 (NOA_DCS_OK) or a negative NOA_DCS_E* code; the message of the
 *     last failure on the calling thread is noa_dcs_last_error();
 *   - all array pointers are device pointers owned by the caller, K and q are
 *     read, outputs are written; n is a 64-bit element count;
 *   - work is enqueued asynchronously on `stream` (0 = the null stream);
 *   - an output pointer that is NULL skips that process.
 * Process ids follow PUMAS order: 0 bremsstrahlung, 1 pair production,
 * 2 photonuclear, 3 ionisation.
 */
#ifndef NOA_DCS_H
#define NOA_DCS_H

#include <stdint.h>

/* The HIP stream handle, declared exactly as hip_runtime_api.h declares it,
 * so that this header needs no ROCm headers (a C11/C++ repeat of an
 * identical typedef is allowed). */
typedef struct ihipStream_t* hipStream_t;

#ifdef __cplusplus
extern "C" {
#endif

#define NOA_DCS_OK 0
#define NOA_DCS_EINVAL (-1)  /* bad argument: NULL input, n < 0, bad element/mass */
#define NOA_DCS_ELAUNCH (-2) /* a HIP call or kernel launch failed */
#define NOA_DCS_ERANGE (-3)  /* a size exceeds a documented limit */
#define NOA_DCS_ESTATE (-4)  /* object used in the wrong state */

/* Replaces noa::pms::AtomicElement (src/noa/pms/physics.hh:39-43); same
 * field order and layout: A in g/mol, I (mean excitation) in GeV, Z. */
typedef struct noa_dcs_element {
    double A;
    double I;
    int32_t Z;
} noa_dcs_element;

/* The four process outputs; any may be NULL. */
typedef struct noa_dcs_outputs_f64 {
    double* brems;
    double* pair;
    double* photo;
    double* ion;
} noa_dcs_outputs_f64;

typedef struct noa_dcs_outputs_f32 {
    float* brems;
    float* pair;
    float* photo;
    float* ion;
} noa_dcs_outputs_f32;

/* Message of the last failed call on this thread ("" if none). */
const char* noa_dcs_last_error(void);

/* Library version string. */
const char* noa_dcs_version(void);

/*
 * FP64 evaluation of the reference arithmetic (within 1e-10 relative of the
 * reference's CPU path, same zero pattern: the closed forms and exits keep
 * its operation order, the quadrature node loops are rearranged to spend
 * fewer divisions and logarithms). Replaces the CPU adaptors dcs::vmap(f) / dcs::pvmap(f)
 * (src/noa/pms/dcs.hh:35-48, 62-75) for f in {bremsstrahlung,
 * pair_production, photonuclear, ionisation} (dcs.hh:134-443) and the CUDA
 * dcs::cuda::vmap_bremsstrahlung (src/noa/pms/dcs.cuh:30-41): one call
 * evaluates every process whose output pointer is non-NULL.
 */
int noa_dcs_eval_f64(const double* K, const double* q, int64_t n, const noa_dcs_element* element,
                     double mass, const noa_dcs_outputs_f64* out, hipStream_t stream);

/*
 * The same FP64 evaluation on the CPU, for host arrays (K, q and the outputs
 * are host pointers; no stream, returns when done). Replaces the reference's
 * CPU adaptors dcs::vmap / dcs::map (serial, src/noa/pms/dcs.hh:35-60) and
 * dcs::pvmap / dcs::pmap (OpenMP, dcs.hh:62-87; loops of
 * src/noa/utils/common.hh:146-183): n_threads = 1 runs serially, 0 uses the
 * OpenMP default team (OMP_NUM_THREADS), k > 1 uses k threads. The physics is
 * the GPU kernels' FP64 code compiled for the host.
 */
int noa_dcs_eval_f64_host(const double* K, const double* q, int64_t n, const noa_dcs_element* element,
                          double mass, const noa_dcs_outputs_f64* out, int32_t n_threads);

/*
 * PUMAS v1.2.1 DCS tabulation (the physics tables PUMAS builds from its DCS
 * models). Replaces compute_dcs_table (src/noa/3rdparty/_pumas/pumas.c:
 * 11693-11787; declared :1270) with its row fill dcs_tabulate_row
 * (:11422-11507), for the default models (pumas.c:190-198): process 0 SSR
 * bremsstrahlung, 1 SSR pair production, 2 DRSS photonuclear.
 *
 * noa_pumas_n_table_dcs(cutoff): the number of grid nodes n_x of a relative
 * cutoff in (0, 1) (pumas.c:1548-1554), or a negative NOA_DCS_E* code
 * (NOA_DCS_ERANGE beyond 192 nodes, i.e. cutoff below ~2e-12).
 *
 * noa_pumas_dcs_table: Z and A (n_elements each) are host arrays (the
 * element invariants are computed on the host with its libm, as the
 * reference computes them); the kinetic energies K (n_K, GeV) are a device
 * array. Outputs (device, caller-owned), in PUMAS's own layouts
 * (pumas.c:4804-4827):
 *   table    [3][n_elements][n_K][2 n_x] floats: per node ln DCS and the
 *            PCHIP slope over the first run of positive DCS values, 0
 *            elsewhere (table_get_dcs);
 *   x        [n_x] floats, the nodes ln(q / K) (table_DCS_x), or NULL;
 *   envelope [3][n_elements][n_K][2] floats, or NULL: entry 0 the envelope
 *            exponent (with the low-energy fill of pumas.c:11745-11761),
 *            entry 1 the envelope maximum, max over the row's kinematic
 *            range above the cutoff of DCS x^exponent, the DCS interpolated
 *            from the table as PUMAS evaluates it (dcs_tabulate_envelope_row,
 *            pumas.c:11526-11679);
 *   table_f64 the table in FP64 before its float rounding, or NULL.
 * csn, the cross-section normalisation table_CSn [3][n_elements][n_K]
 * (device doubles), gates rows as PUMAS does: a row whose value is not
 * positive is all zeros and has no envelope (pumas.c:11427-11434, 11532-11536;
 * its exponent is written as 0). NULL: every row positive. Work is enqueued
 * on `stream`.
 *
 * noa_pumas_dcs_f64: the model DCS of one process (m^2/kg, the wrappers of
 * pumas.c:11067-11112) over n device (K, q) pairs.
 */
int noa_pumas_n_table_dcs(double cutoff);
int noa_pumas_dcs_table(const double* Z, const double* A, int32_t n_elements, double mass, double cutoff,
                        const double* K, int32_t n_K, float* table, float* x, float* envelope, double* table_f64,
                        const double* csn, hipStream_t stream);
int noa_pumas_dcs_f64(int32_t process, double Z, double A, double mass, const double* K, const double* q, int64_t n,
                      double* out, hipStream_t stream);

/*
 * Mixed-precision evaluation: FP32 arithmetic with FP64 kept for the
 * kinematic differences (E - q, K - q, 1 - nu), the pair-production
 * integration bound and the photonuclear Q^2 range. FP64 inputs, FP32
 * outputs. No reference counterpart (the reference is FP64 only); same call
 * shape as noa_dcs_eval_f64.
 */
int noa_dcs_eval_f32mixed(const double* K, const double* q, int64_t n, const noa_dcs_element* element,
                          double mass, const noa_dcs_outputs_f32* out, hipStream_t stream);

/*
 * DEL and CEL recoil integrals of the four processes for n kinetic energies.
 * Replaces dcs::vmap_integral(dcs::recoil_integral(f, dcs::del_integrand))
 * and the cel_integrand form (src/noa/pms/dcs.hh:89-130), including the
 * analytic ionisation branch below K = (m - m_e)^2 / (2 m_e)
 * (dcs.hh:444-497, 955-1001):
 *   del[p][i] = (1/E) int_{K xlow}^{K} dcs_p(K, q) dq,   cel[p][i] = (1/E) int dcs_p(K, q) q dq,
 * integrated over t = ln q (integrand dcs q, resp. dcs q^2) with the 6-node
 * rule on ceil(min_points/6) subintervals.
 * Any output pointer may be NULL (del or cel themselves too). workspace is
 * a device buffer of at least noa_dcs_recoil_workspace_bytes(n, min_points).
 */
int64_t noa_dcs_recoil_workspace_bytes(int64_t n, int32_t min_points);
int noa_dcs_recoil_integrals_f64(const double* K, int64_t n, double xlow, const noa_dcs_element* element,
                                 double mass, int32_t min_points, const noa_dcs_outputs_f64* del,
                                 const noa_dcs_outputs_f64* cel, void* workspace, int64_t workspace_bytes,
                                 hipStream_t stream);

/*
 * The same recoil integrals with no allocation and no workspace argument.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). It calls no hipMalloc / hipFree and no allocating library
 * routine; its node pairs and DCS planes live in a file-scope device array
 * of noa_dcs_recoil_noalloc_scratch_bytes() (32 MiB) that each GPU holds
 * until the process ends. Any n is processed, in chunks of
 * scratch / (48 ceil(min_points / 6) * 6) kinetic energies (3883 at
 * min_points = 180); a min_points whose single kinetic energy would not fit
 * returns NOA_DCS_ERANGE. Results equal noa_dcs_recoil_integrals_f64's bit
 * for bit. Every call on a GPU shares that GPU's array: calls that overlap
 * on one GPU (two streams, or two host threads) race; serialise them.
 */
int64_t noa_dcs_recoil_noalloc_scratch_bytes(void);
int noa_dcs_recoil_integrals_f64_noalloc(const double* K, int64_t n, double xlow, const noa_dcs_element* element,
                                         double mass, int32_t min_points, const noa_dcs_outputs_f64* del,
                                         const noa_dcs_outputs_f64* cel, hipStream_t stream);

/* ---- Coulomb scattering (src/noa/pms/dcs.hh:499-950) ----
 * Arrays are row-major like the reference's tensors: fCM [n][2],
 * screening [n][9], G [n][2]; for a mixture of nel elements the per-element
 * arrays are [nel][nkin][...] (element-major).
 *
 * Replaces dcs::coulomb_data (dcs.hh:604-622): per kinetic energy, the
 * centre-of-mass factors, the spin factor, the 3 screening parameters and 6
 * pole-reduction coefficients, and the inverse Wentzel path.
 */
int noa_dcs_coulomb_data_f64(const double* K, int64_t n, const noa_dcs_element* element, double mass, double* fCM,
                             double* screening, double* fspin, double* invlambda, hipStream_t stream);
/* Replaces dcs::coulomb_transport (dcs.hh:676-694): first and second transport
 * coefficients G at the cutoff mu (n_mu = 1: one mu for all; n_mu = n: per K). */
int noa_dcs_coulomb_transport_f64(const double* screening, const double* fspin, const double* mu, int64_t n_mu,
                                  int64_t n, double* G, hipStream_t stream);
/* Replaces dcs::hard_scattering (dcs.hh:842-872): per kinetic energy, the
 * hard-scattering cutoff angle mu0 (Ridders' root, numerics.hh:154-214) and
 * mean free path lb_h of a mixture of nel elements. */
int noa_dcs_hard_scattering_f64(int64_t nel, int64_t nkin, const double* G, const double* fCM,
                                const double* screening, const double* invlambda, const double* fspin, double* mu0,
                                double* lb_h, hipStream_t stream);
/* Replaces dcs::soft_scattering (dcs.hh:940-950): the first transverse
 * transport coefficient of soft ionisation and photonuclear interactions.
 * workspace is a device buffer of noa_dcs_soft_scattering_workspace_bytes(n). */
int64_t noa_dcs_soft_scattering_workspace_bytes(int64_t n);
int noa_dcs_soft_scattering_f64(const double* K, int64_t n, const noa_dcs_element* element, double mass,
                                double* result, void* workspace, int64_t workspace_bytes, hipStream_t stream);

/* ---- batched uneven elements: one persistent launch, captured in a hipGraph ----
 * Element e owns pairs [offsets[e], offsets[e+1]) of the concatenated K, q
 * and output arrays. The reference has no batched form: a caller loops
 * dcs::vmap over elements (src/noa/pms/dcs.hh:35-48). The plan binds the
 * buffers given at creation; noa_dcs_batch_launch replays the captured graph.
 */
typedef struct noa_dcs_batch noa_dcs_batch;

#define NOA_DCS_PREC_F64 0
#define NOA_DCS_PREC_F32MIXED 1

typedef struct noa_dcs_batch_config {
    int32_t precision;      /* NOA_DCS_PREC_F64 or NOA_DCS_PREC_F32MIXED */
    int32_t tile;           /* pairs per work item; 0 = autotuned */
    int32_t waves_per_cu;   /* persistent workgroups per CU; 0 = autotuned */
    int32_t use_graph;      /* 1 = capture the launch in a hipGraph */
    const char* tune_cache; /* autotune cache file, NULL = none */
    int32_t unroll;         /* unroll of the integrals' node loops, 1 or 2; 0 = autotuned */
} noa_dcs_batch_config;

/* elements and offsets are HOST arrays (n_elements and n_elements + 1
 * entries); K, q and the outputs are device arrays of offsets[n_elements]
 * pairs (outputs as double* or float* per precision, cast to void*). */
int noa_dcs_batch_create(const noa_dcs_element* elements, const int64_t* offsets, int32_t n_elements,
                         double mass, const double* K, const double* q, void* out_brems, void* out_pair,
                         void* out_photo, void* out_ion, const noa_dcs_batch_config* config,
                         hipStream_t stream, noa_dcs_batch** plan);
int noa_dcs_batch_launch(noa_dcs_batch* plan, hipStream_t stream);
/* Chosen tile, workgroups per CU and node-loop unroll, and the number of
 * work items; any output pointer may be NULL. */
int noa_dcs_batch_info(const noa_dcs_batch* plan, int32_t* tile, int32_t* waves_per_cu, int32_t* unroll,
                       int64_t* n_items);
int noa_dcs_batch_destroy(noa_dcs_batch* plan);

/* ---- learned surrogate of the four DCS processes ----
 * No reference counterpart: the reference evaluates only the exact formulas.
 * An MLP (log10 K, log10 nu, Z) -> 4 hidden layers of 128 with SiLU -> the
 * standardised ln(A * DCS) of the 4 processes; hidden and output layers on
 * bf16 MFMA with FP32 accumulation, the first layer in FP32. Parameters are
 * a flat FP32 vector of noa_dcs_mlp_num_params() values (W[out][in]
 * row-major per layer, then its bias; layers 1..5). X is [n x 3] FP32
 * features, T and M are [n x 4] FP32 targets and 0/1 masks (device). */
typedef struct noa_dcs_mlp noa_dcs_mlp;

int64_t noa_dcs_mlp_num_params(void);
/* max_rows bounds the training batch and the forward chunk (<= 2^20). */
int noa_dcs_mlp_create(uint64_t seed, int32_t max_rows, noa_dcs_mlp** mlp);
int noa_dcs_mlp_destroy(noa_dcs_mlp* mlp);
int noa_dcs_mlp_get_params(const noa_dcs_mlp* mlp, float* host, int64_t n);
int noa_dcs_mlp_set_params(noa_dcs_mlp* mlp, const float* host, int64_t n);
/* Features of n pairs of one element: x = ((log10 K - 1.5) / 4.5,
 * (log10(q/K) + 3) / 3, Z / 46 - 1). */
int noa_dcs_mlp_features(const double* K, const double* q, int64_t n, int32_t Z, float* X, hipStream_t stream);
/* Layer-by-layer forward with the training kernels: Y [rows x 4]. */
int noa_dcs_mlp_forward(noa_dcs_mlp* mlp, const float* X, int64_t rows, float* Y, hipStream_t stream);
/* One forward + backward of the masked mean squared error on rows <=
 * max_rows; loss and the gradient (num_params floats) to host memory. */
int noa_dcs_mlp_forward_backward(noa_dcs_mlp* mlp, const float* X, const float* T, const float* M, int32_t rows,
                                 double* loss, float* grads_host, hipStream_t stream);
/* Adam over minibatches of a device dataset of n rows for `epochs` epochs,
 * cosine learning rate lr_max -> lr_min; loss_per_epoch: host, may be NULL. */
int noa_dcs_mlp_train(noa_dcs_mlp* mlp, const float* X, const float* T, const float* M, int64_t n, int32_t batch,
                      int32_t epochs, float lr_max, float lr_min, uint64_t seed, float* loss_per_epoch,
                      hipStream_t stream);
/* The fused inference kernel (all layers chained in registers): Y [n x 4]. */
int noa_dcs_mlp_infer_fused(noa_dcs_mlp* mlp, const float* X, int64_t n, float* Y, hipStream_t stream);

typedef struct noa_dcs_surrogate_cfg {
    float out_mean[4];      /* de-standardisation of the outputs: ln(A * DCS_p) = */
    float out_std[4];       /*   out_mean[p] + sum_k trend[p][k] phi_k(x) + out_std[p] * y_p with the */
    float trend[4][13];     /*   cubic basis phi(x) = (a, b, a^2, b^2, ab, a^3, b^3, a^2 b, a b^2, z, za, zb,
                                 z^2) of the features x = (a, b, z) */
    const uint8_t* flag_map; /* device [n_elements][bins][bins] over (log10 K in [-3,6], log10 nu in [-6,0]);
                                bit p set = evaluate process p exactly in that cell; may be NULL */
    int32_t bins;
    int32_t recheck_per_mille; /* seeded sample evaluated exactly and compared (10 = 1%) */
    uint64_t seed;
} noa_dcs_surrogate_cfg;

/* Surrogate DCS of the queries of an element batch (layout as in
 * noa_dcs_batch_create), FP32 outputs. Queries outside the training domain
 * and the seeded recheck sample take every process from the exact FP64
 * physics, other queries the processes flagged in their cell; the exact
 * evaluations run in the same kernel. counters is a device array of 4
 * uint32: on completion counters[1] and counters[3] = number of
 * pair-production and photonuclear evaluations done exactly, counters[2] =
 * bits of the largest relative error of the network over the recheck sample
 * (FP32, processes the network serves in that cell). Asynchronous on
 * `stream`; calls on one network are ordered by the stream. */
int noa_dcs_surrogate_eval(noa_dcs_mlp* mlp, const noa_dcs_element* elements, const int64_t* offsets,
                           int32_t n_elements, double mass, const double* K, const double* q,
                           const noa_dcs_outputs_f32* out, const noa_dcs_surrogate_cfg* cfg, uint32_t* counters,
                           hipStream_t stream);

/* ---- synthetic workload (counter-based, seeded) ----
 * u(seed, stream, i) is a 53-bit uniform in [0, 1) from a SplitMix64 output
 * of a (seed, stream, i) counter. Fills K[i] = 10^(lk0 + (lk1 - lk0) u0),
 * nu = 10^(ln0 + (ln1 - ln0) u1), q[i] = nu K[i], with u0 = u(seed, 2s, i),
 * u1 = u(seed, 2s + 1, i), s = stream. */
int noa_dcs_fill_workload(double* K, double* q, int64_t n, uint64_t seed, uint64_t stream_id, double lk0,
                          double lk1, double ln0, double ln1, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* NOA_DCS_H */
