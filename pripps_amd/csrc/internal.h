// Internal layouts shared by the HIP kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

namespace pripps {

// Maximum number of distinct atom kinds in one structure; classes are
// (kind, hydrated) pairs, so there are at most 2 * kMaxKinds of them.
constexpr int kMaxKinds = 64;
constexpr int kMaxClasses = 2 * kMaxKinds;
constexpr int kProfiles = 6;  // DebyeTerms order: AA EE HH AE AH EH
constexpr int kFinChunk = 32; // item partials summed per block in the first reduction stage
// q-points per strip of the Chebyshev pair kernel (debye_cheb.hip); strips of
// that kernel are runs of up to kChebQ consecutive q-points with their own
// centre and half-width, on any grid.
constexpr int kChebQ = 256;
// FP64 coefficient sums per class run of a Chebyshev item: S = sum s,
// Si = sum w_i s (pass 0), Sw = sum w_j s, Siw = sum w_i w_j s (pass 1,
// hydrated j)
constexpr int kChebSets = 4;
// expansion orders the Chebyshev kernel is built for (ascending)
constexpr int kChebOrders[] = {12, 16, 20, 22, 24, 26, 28, 30, 32, 36, 40};
constexpr int kChebNumOrders = sizeof(kChebOrders) / sizeof(kChebOrders[0]);
// ... and its FP64 form, whose orders keep the truncation below 1e-15
constexpr int kCheb64Orders[] = {16, 20, 24, 28, 32, 36};
constexpr int kCheb64NumOrders = sizeof(kCheb64Orders) / sizeof(kCheb64Orders[0]);

// One atom of the sorted SoA-in-AoS stream: centred coordinates and the
// hydration weight, 16 bytes (FP32) or 32 bytes (FP64) per atom so a wave
// fetches one j atom with a single scalar load.
template <typename T>
struct alignas(4 * sizeof(T)) Atom {
    T x, y, z, w;
};

// Per-structure description, resident in device memory.
struct DevStruct {
    const void *atoms;      // Atom<T>[n] sorted by class
    const double *q;        // [nq] q values
    const double *ffa;      // [K][nq] atom form factors
    const double *ffe;      // [K][nq] excluded-volume form factors or nullptr
    const double *ffw;      // [nq] water form factor or nullptr
    double *profiles;       // [6][nq] output
    const double *kind_n;   // [K] atoms per kind
    const double *kind_w;   // [K] sum of hydration weights per kind
    double sum_w2;          // sum of squared hydration weights
    double q0, dq;          // uniform grid: q_m = q0 + m dq
    int32_t n, nq, K, uniform;
    int32_t small_h;        // dq r_max / 4 pi <= 1/8 revolution for every pair
    int32_t n_classes;
    int32_t n_strips;       // ceil(nq / Q)
    int32_t block;          // j atoms summed in FP32 before the FP64 carry
    int32_t sin_direct;     // q_max r_max / 2 pi < 255 revolutions for every pair: the
                            // hardware sine takes the phase without a fract
    // Chebyshev plans only (else nullptr): strip s covers q-points
    // [strip_first[s], strip_first[s + 1]) with centre strip_geo[2 s] and
    // half-width strip_geo[2 s + 1]
    const int32_t *strip_first;
    const double *strip_geo;
    int32_t cls_start[kMaxClasses + 1];
    int8_t cls_kind[kMaxClasses];
    int8_t cls_wet[kMaxClasses];
};

// One wave's work: a 64-atom i-tile of one kind against a j-span, for one
// q strip of Q points.
struct Item {
    int32_t s;       // structure index
    int32_t i0;      // first (sorted) atom of the i-tile
    int32_t icount;  // atoms in the tile, 1..64
    int32_t ikind;   // kind of every atom in the tile
    int32_t j0, j1;  // j span [j0, j1), j0 >= i0
    int32_t strip;   // q strip index
    int32_t pad;
};

// Chebyshev plans: one wave's share of a structure's work for one strip
// group -- a contiguous run of items, cut so that the units of a structure
// cost about the same -- summed into one partial slot per strip.
struct Unit {
    int32_t item0, item1;  // items [item0, item1)
    int32_t rec0;          // record of its first class run (one per run, in order)
    int32_t nk;            // strips in the group (1 or 2)
};

// Reduction of a Chebyshev launch.  A record holds the FP64 coefficient sums
// of one class run: [kChebSets][RS] (set, then coefficient k P + p of strip k
// of the group; RS = 64 or 128).  A combo is (structure, strip group, i kind,
// j kind); its records are summed in chunks, the chunk totals of a combo are
// summed and evaluated at the q-points of each strip of the group (stage 2,
// one block per (strip, combo)), and stage 3 adds a strip's combos, divides
// by q and adds the diagonal.
struct ChebReduceArgs {
    const DevStruct *structs;
    const double *records;       // [records][kChebSets][RS]
    int RS, P;
    const int32_t *chunk_recs;   // record ids, combo by combo, in queue order
    const int32_t *chunk_first;  // [n_chunks + 1] into chunk_recs
    int n_chunks;
    double *chunk_tot;           // [n_chunks][kChebSets][RS]
    const int4 *sb_a;            // stage-2 block: (structure, strip, k in group, i kind)
    const int4 *sb_b;            // (j kind, first chunk, end chunk, 0)
    int n_sb;
    double *sb_out;              // [n_sb][6][kChebQ]
    const int4 *st;              // stage-3 block: (structure, strip, first sb, end sb)
    int n_st;
};

// Start values of the Chebyshev kernel's Bessel recurrence for one order P
// (host-computed, cheb_limits): z = (beta / 2)^2 is clamped to z_floor, the
// start at order P + 2 is clamp(fma(z, k1, b1), 0, 2^-113) and, for P > 14,
// the restart at order 14 is clamp(fma(z, k2, b2), 0, 2^-126).
struct ChebLimits {
    float z_floor, k1, b1, k2, b2;
};

// The same for the FP64 Chebyshev kernel (cheb64_limits): z clamped to
// z_floor, the start at order P + 2 is clamp(fma(z, k1, b1), 0, 2^-600).
struct ChebLimits64 {
    double z_floor, k1, b1;
};

// Kernel argument blocks.
struct TiledArgs {
    const DevStruct *structs;
    const Item *items;
    int64_t n_items;   // queue length: items, or units (Chebyshev plans)
    int *counter;      // work-queue head (and, Chebyshev plans, exit count): zero at launch
    double *partials;  // [slots][6][Q]; a slot per item, or per unit and strip
    int variant;       // FP32 uniform inner-loop variant (see debye_tiled_kernel)
    const Unit *units; // Chebyshev plans, else nullptr
    ChebLimits cheb;   // Chebyshev plans: set by launch_cheb for the launch's order
    ChebLimits64 cheb64;
};

struct FinalizeArgs {
    const DevStruct *structs;
    const int32_t *csr_ptr;       // per (structure, strip): offsets into csr_items
    const int32_t *csr_items;     // item ids
    const double *partials;
    const int32_t *strip_struct;  // flattened (structure, strip) -> structure
    const int32_t *strip_index;   // flattened (structure, strip) -> strip
    const int32_t *stage_ptr;     // per (structure, strip): first chunk slot in `stage`
    double *stage;                // [chunks][6][Q] chunk totals
    int n_strips;                 // (structure, strip) pairs
    int total_chunks;             // stage_ptr[n_strips]: one reduce block each
    int Q;
};

struct PlainArgs {
    const DevStruct *sd;
    const int4 *tiles;  // (i0, icount, ikind, 0)
    int n_tiles;
    double *partials;   // [n_tiles][nq][6]
};

// What another translation unit needs of an engine (engine.hip).
struct EngineAccess {
    int device;
    int n_cu;
    hipStream_t stream;
    hipEvent_t ev0, ev1;  // timing events, used under `mu`
    std::mutex *mu;       // serialises submissions on `stream`
};

// launch wrappers (debye_kernels.hip)
hipError_t launch_tiled(bool fp64, int Q, bool uniform, const TiledArgs &a, int blocks,
                        hipStream_t st);
hipError_t tiled_occupancy(bool fp64, int Q, bool uniform, int variant, int *blocks_per_cu);
hipError_t launch_finalize(const FinalizeArgs &a, int n_strips_total, hipStream_t st);
// Chebyshev pair kernel (debye_cheb.hip), P = expansion order (a built order)
hipError_t launch_cheb(int P, const TiledArgs &a, int blocks, hipStream_t st);
hipError_t cheb_occupancy(int P, int *blocks_per_cu);
// start values of the pair recurrence for order P; false if P is not supported
bool cheb_limits(int P, ChebLimits *out);
// FP64 Chebyshev pair kernel, its occupancy, start values and truncation limit
// (1e-15 of a pair's amplitude)
hipError_t launch_cheb64(int P, const TiledArgs &a, int blocks, hipStream_t st);
hipError_t cheb64_occupancy(int P, int *blocks_per_cu);
bool cheb64_limits(int P, ChebLimits64 *out);
double cheb64_beta_limit(int P);
hipError_t launch_cheb_reduce(const ChebReduceArgs &a, hipStream_t st);
// smallest built order whose truncation is below 1e-8 of a pair's amplitude
// at beta = H r up to beta_max (0: none)
int cheb_order_for(double beta_max);
// the largest beta_max that order P admits
double cheb_beta_limit(int P);
hipError_t launch_plain(const PlainArgs &a, int nq, hipStream_t st);
hipError_t launch_fill_int(int *p, int n, int v, hipStream_t st);
hipError_t launch_combine(const double *prof, int nq, double vs, double cd, double *out,
                          hipStream_t st);
hipError_t launch_fill_double(double *p, int64_t n, double v, hipStream_t st);
// acc[6 nq] += sum_c w[c] prof[c][6 nq], c = 0..n_slots-1 in order
hipError_t launch_frames_accumulate(const double *prof, int n_slots, const double *w, int nq,
                                    double *acc, hipStream_t st);
// out[6 nq] = acc[6 nq] / wsum
hipError_t launch_frames_normalize(const double *acc, int nq, double wsum, double *out,
                                   hipStream_t st);

}  // namespace pripps
