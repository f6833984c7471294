// One tile of (K, q) pairs through all four DCS processes, with in-LDS
// compaction of the two quadrature processes.
//
// Pair production is active on ~59% and photonuclear on ~23% of uniformly
// drawn pairs, so a lane-per-pair map leaves most of a wave idle inside the
// 8- and 9-node integrals. A tile is therefore evaluated in phases:
//   1. each lane takes TILE/WG pairs with coalesced loads, computes the two
//      closed-form processes (bremsstrahlung, ionisation), runs the cheap
//      kinematic exits of the two integral processes, writes 0 where they
//      exit, and appends the surviving pair to a per-process list in LDS
//      (wave ballot + one LDS atomic per wave; the FP64 form assembles the
//      pair-production plane in LDS, starting from the argmin of its exit
//      test, and writes it out coalesced after phase 2);
//   2. all lanes drain the two lists densely, heaviest process first, so
//      every wave runs the integrals with (almost) all lanes active;
//   3. (mixed precision only) pairs outside the FP32 form's range were
//      appended to a third list in phase 1 (all four processes), and pairs
//      whose FP32 integral failed its own error estimate to the same list in
//      phase 2 (that integral); they are recomputed densely with the FP64
//      physics.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "dcs_consts.hh"
#include "dcs_f64.hh"
#include "fastmath.hh"

namespace noa_hip {

constexpr int WG = 256;  // 4 wave64s

template <class OutT>
struct Outs {
    OutT* b;
    OutT* p;
    OutT* ph;
    OutT* i;
};

// The tile's lists in LDS. A pair enters the redo list at most twice (once
// for all four processes from phase 1, or once per failed integral from
// phase 2), so a form with a redo list holds 2 TILE entries; the FP64 form
// holds none. The FP64 form assembles the tile's pair-production plane in
// LDS (pair_out, by tile offset): the lower end of the integral (argmin),
// which the exit test of phase 1 has already computed, until phase 2
// replaces it with the result, and the exit's 0 where the integral does not
// run; the plane is then written out with coalesced stores.
template <int TILE, bool REDO>
struct TileLists {
    static_assert(TILE <= 16384, "14-bit tile offsets (bits 14 and 15 tag the process in the redo list)");
    double pair_out[REDO ? 1 : TILE];
    uint16_t photo[TILE];
    uint16_t pair[TILE];
    uint16_t redo[REDO ? 2 * TILE : 1];
    uint32_t n_photo;
    uint32_t n_pair;
    uint32_t n_redo;
};

constexpr uint32_t REDO_PHOTO = 0x8000u;  // the photonuclear integral of the pair
constexpr uint32_t REDO_ALL = 0x4000u;    // all four processes of the pair
constexpr uint32_t REDO_OFFSET = 0x3FFFu;

// Appends `value` for every active lane with pred set; one LDS atomic per
// wave. Lanes outside the current EXEC mask take no part, so it may be called
// from divergent code as long as every lane with pred set is active. Returns
// the entry's position in the list (meaningful where pred is set).
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t* counter, uint16_t* list, uint32_t value) {
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return 0u;
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
    uint32_t base = 0;
    if (pred && rank == 0) base = atomicAdd(counter, uint32_t(__popcll(m)));
    const int leader = __ffsll((long long)m) - 1;
    base = uint32_t(__builtin_amdgcn_readlane(int(base), leader));
    if (pred) list[base + rank] = uint16_t(value);
    return base + rank;
}

// Output stores are non-temporal: the results are not read again by the
// kernel, and keeping them out of L2's retained lines leaves room for the
// tile's K and q, which phase 2 re-reads (measured on 2^24 pairs: HBM fetch
// 47 -> 38 B per pair, 0.6% faster). Staging K and q in LDS instead would
// bring the fetch to the algorithmic 16 B per pair, but its 32 KiB per tile
// cut the occupancy from 5 to 3 waves per SIMD and the kernel 8% slower.
template <class T>
__device__ __forceinline__ void store_stream(T* p, T v) {
    __builtin_nontemporal_store(v, p);
}

// The launch constants, addressed in the constant address space (the kernel
// argument segment, or the batch's element table): read with scalar loads.
using ConstsPtr = const __attribute__((address_space(4))) DcsConsts*;

// The constants for one item of a phase loop. The empty asm hides the
// pointer's value, so the compiler re-reads the fields each item (scalar
// loads, which cost no VALU slot) instead of keeping ~48 of them in SGPRs
// across the whole tile: held that long, they overflow the SGPR file and
// are spilled to VGPR lanes, and every use then costs a v_readlane (measured
// on 2^24 pairs: 1.7% off the FP64 kernel, 5% off the mixed form, whose
// spills were the heavier).
__device__ __forceinline__ const DcsConsts& item_consts(ConstsPtr p) {
    asm volatile("" : "+s"(p));
    return *(const DcsConsts*)p;
}

// The leading arguments of a kernel that takes (K, q, n, DcsConsts c, ...):
// the kernel-argument segment holds them in declaration order at their
// natural alignment, as this struct lays them out.
struct TileKernelArgs {
    const double* K;
    const double* q;
    int64_t n;
    DcsConsts c;
};

// c of such a kernel, read in place from the kernel-argument segment (the
// address of the parameter as a C++ object is a private copy of all its
// bytes in scratch).
__device__ __forceinline__ ConstsPtr kernarg_consts() {
    using KArg = const __attribute__((address_space(4))) char*;
    return (ConstsPtr)((KArg)__builtin_amdgcn_kernarg_segment_ptr() + offsetof(TileKernelArgs, c));
}

// Copies the log/exp tables into an LDS block (visible after the next barrier).
__device__ __forceinline__ void load_fm_tables(double* block) {
    for (int k = threadIdx.x; k < fm::TABLE_DOUBLES; k += WG) block[k] = fm::table_entry(k);
}

// Phase 1 of run_tile: each lane takes TILE / WG pairs with coalesced loads,
// evaluates the closed forms and the exits of the two integrals, and appends
// the pairs the integrals need to the tile's lists. ALL: every output is
// requested (no per-output branches).
template <class P, int TILE, bool ALL>
__device__ __forceinline__ void phase1(int64_t base, int64_t end, const double* __restrict__ K,
                                       const double* __restrict__ q, ConstsPtr cp,
                                       const Outs<typename P::out_t>& o, TileLists<TILE, P::has_redo>& s,
                                       const fm::Tab& t) {
    using out_t = typename P::out_t;
    constexpr int ITEMS = TILE / WG;
    const int tid = threadIdx.x;
    const bool want_pair = ALL || o.p != nullptr;
    const bool want_photo = ALL || o.ph != nullptr;
    int64_t gi = base + tid;
    double k_next = 0., q_next = 0.;
    if (gi < end) {
        k_next = K[gi];
        q_next = q[gi];
    }
#pragma unroll 1
    for (int it = 0; it < ITEMS; it++) {
        const DcsConsts& c = item_consts(cp);
        const int local = it * WG + tid;
        gi = base + local;
        const bool valid = gi < end;
        const double k = k_next, qq = q_next;
        if (it + 1 < ITEMS && gi + WG < end) {  // prefetch the next pair of this lane
            k_next = K[gi + WG];
            q_next = q[gi + WG];
        }
        bool pa = false, pha = false, full_redo = false;
        double am = 0.;  // (FP64 form) argmin of an active pair-production integral
        // (mixed form) a pair outside the FP32 range goes whole to phase 3
        if constexpr (P::has_redo) full_redo = valid && !P::in_range(k, qq, c);
        if (valid && !full_redo) {
            if constexpr (ALL) {
                const out_t b = P::brems(k, qq, c, t);
                const out_t i = P::ion(k, qq, c, t);
                pa = P::pair_active(k, qq, c, am);
                pha = P::photo_active(k, qq, c);
                store_stream(&o.b[gi], b);
                store_stream(&o.i[gi], i);
                if (!pha) store_stream(&o.ph[gi], out_t(0));
            } else {
                if (o.b) store_stream(&o.b[gi], P::brems(k, qq, c, t));
                if (o.i) store_stream(&o.i[gi], P::ion(k, qq, c, t));
                if (want_pair) {
                    if constexpr (P::has_redo) {
                        pa = P::pair_active(k, qq, c);
                        if (!pa) store_stream(&o.p[gi], out_t(0));
                    } else {
                        pa = P::pair_active(k, qq, c, am);
                    }
                }
                if (want_photo) {
                    pha = P::photo_active(k, qq, c);
                    if (!pha) store_stream(&o.ph[gi], out_t(0));
                }
            }
        }
        if (want_photo) wave_append(pha, &s.n_photo, s.photo, uint32_t(local));
        if (want_pair) {
            wave_append(pa, &s.n_pair, s.pair, uint32_t(local));
            // (FP64 form) the pair-production plane of the tile is assembled in
            // LDS: argmin where the integral runs (replaced by its result in
            // phase 2), the exit's 0 elsewhere; written out coalesced at the end
            if constexpr (!P::has_redo)
                if (valid) s.pair_out[local] = pa ? am : 0.;
        }
        if constexpr (P::has_redo) wave_append(full_redo, &s.n_redo, s.redo, uint32_t(local) | REDO_ALL);
    }
}

// Evaluates pairs [base, end) (end - base <= TILE) of one element with the
// log/exp tables `t` in LDS. Must be called by all WG threads of the
// workgroup.
template <class P, int TILE>
__device__ __forceinline__ void run_tile(int64_t base, int64_t end, const double* __restrict__ K,
                                         const double* __restrict__ q, ConstsPtr cp,
                                         const Outs<typename P::out_t>& o, TileLists<TILE, P::has_redo>& s,
                                         const fm::Tab& t) {
    using out_t = typename P::out_t;
    const int tid = threadIdx.x;
    if (tid == 0) {
        s.n_photo = 0;
        s.n_pair = 0;
        s.n_redo = 0;
    }
    __syncthreads();

    // phase 1: closed forms, exits, compaction. In the FP64 form with all
    // four outputs requested (the usual call) the four evaluations of a pair
    // sit in one block, so the compiler shares what they have in common (the
    // fast-path test, E = K + m and its reciprocal, 1 / q): 1.4% on the
    // kernel. The mixed form keeps one version (a second raised its register
    // spills and cost it 0.6%).
    if constexpr (!P::has_redo) {
        if (o.b && o.i && o.p && o.ph) {
            phase1<P, TILE, true>(base, end, K, q, cp, o, s, t);
        } else {
            phase1<P, TILE, false>(base, end, K, q, cp, o, s, t);
        }
    } else {
        phase1<P, TILE, false>(base, end, K, q, cp, o, s, t);
    }
    __syncthreads();

    // phase 2: dense integrals, photonuclear (heavier) first
    const uint32_t nph = s.n_photo;
    const uint32_t total = nph + s.n_pair;
#pragma unroll 1
    for (uint32_t v = tid; v < total; v += WG) {
        const DcsConsts& c = item_consts(cp);
        bool redo = false;
        uint32_t tag;
        if (v < nph) {
            tag = s.photo[v];
            const int64_t g = base + tag;
            store_stream(&o.ph[g], P::photo(K[g], q[g], c, redo, t));
            tag |= REDO_PHOTO;
        } else {
            tag = s.pair[v - nph];
            const int64_t g = base + tag;
            if constexpr (P::has_redo)
                store_stream(&o.p[g], P::pair(K[g], q[g], c, redo, t));
            else
                s.pair_out[tag] = P::pair_at(K[g], q[g], s.pair_out[tag], c, t);
        }
        if (P::has_redo) wave_append(redo, &s.n_redo, s.redo, tag);
    }

    // (FP64 form) the assembled pair-production plane, coalesced
    if constexpr (!P::has_redo) {
        if (o.p) {
            __syncthreads();
            for (int64_t gi = base + tid; gi < end; gi += WG) store_stream(&o.p[gi], s.pair_out[gi - base]);
        }
    }

    // phase 3: FP64 recomputation of the pairs that failed their estimate
    if (P::has_redo) {
        __syncthreads();
        const uint32_t nr = s.n_redo;
#pragma unroll 1
        for (uint32_t v = tid; v < nr; v += WG) {
            const DcsConsts& c = item_consts(cp);
            const uint32_t tag = s.redo[v];
            const int64_t g = base + (tag & REDO_OFFSET);
            if (tag & REDO_ALL) {
                const double k = K[g], qq = q[g];
                if (o.b) o.b[g] = out_t(F64::brems(k, qq, c, t));
                if (o.i) o.i[g] = out_t(F64::ion_redo(k, qq, c, t));
                if (o.p) o.p[g] = out_t(F64::pair_active(k, qq, c) ? F64::pair_ieee(k, qq, c, t) : 0.);
                if (o.ph) o.ph[g] = out_t(F64::photo_active(k, qq, c) ? F64::photo_ieee(k, qq, c, t) : 0.);
            } else if (tag & REDO_PHOTO) {
                o.ph[g] = out_t(F64::photo_ieee(K[g], q[g], c, t));
            } else {
                o.p[g] = out_t(F64::pair_ieee(K[g], q[g], c, t));
            }
        }
    }
    __syncthreads();  // the lists are reused by the next tile of a persistent workgroup
}

}  // namespace noa_hip
