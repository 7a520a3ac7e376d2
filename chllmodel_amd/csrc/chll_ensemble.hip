// chll_ensemble.hip -- many replicas of uneven size and (kT, g) swept by one
// launch per half-sweep, with parallel-tempering swaps on the device, captured
// in a hipGraph.
//
// The reference scans temperature / field grids by running one ChLLSim after
// another (e.g. scripts/phase_diagram_large1.py loops over K x T); here every
// replica of a scan lives in HBM at once.  A launch's block b belongs to one
// replica (table blk_rep, replica-major, largest replicas first) and runs the
// same per-site trial as the single-lattice kernels (chll_site.h).
//
// Replicas are grouped into ladders: replicas of one ladder share (ni, nj, nk,
// g, K) and sit at distinct temperature slots, listed in ascending kT.  Each
// slot owns its kT and its adaptive cone step d; a parallel-tempering swap
// exchanges the slots of two replicas at neighbouring temperatures with
// probability min(1, exp((1/kT_t - 1/kT_{t+1}) (E_a - E_b))), E the replica's
// binary64 lattice Hamiltonian.
//
// One "block" of K sweeps is: K x (four half-sweeps + a per-replica epilogue),
// then the observable reduction, the measuring epilogue and the swap kernel,
// which also advances the device sweep counter.  Nothing reads host memory, so
// the block is captured once and replayed.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

#include "../../include/chll_mod.h"
#include "chll_site.h"

namespace chll {

constexpr int kEnsCols = CHLL_ENS_COLS;

// Replica descriptor in device memory (read by whole blocks: uniform loads).
struct RepDesc {
    Geom G;
    Fields F;
    uint32_t blk_start, nblk;  // this replica's blocks in a launch
    uint32_t replica;          // Philox counter word
    uint32_t nmeas;            // its measurement partials (blocks of the measure launch)
    double g, KK;
    uint32_t *counts;          // kGroups x kSlots x kSlotStride (groups as the single lattice)
    // reference arithmetic fused through held trials (chll_site.h), else null:
    // colour 1's flip bitmap of the current sweep and its held-trial queue
    unsigned long long *flipbits;
    uint32_t *held;
};

// Ensemble-wide device state.
struct EnsDev {
    const RepDesc *R;
    const uint32_t *blk_rep;  // per launch block: replica << 24 | local block
    const uint2 *blk_rep2;    // pair kernels: (replica << 24 | local block, live threads), or null
    const uint32_t *blk_rep3; // k-marching observables: replica << 24 | local block, or null
    const uint32_t *order;    // replicas in block-table order (largest first)
    double *kbt, *inv_kbt;    // per temperature slot
    double *d;                // per temperature slot (adaptive cone step)
    int32_t *rep_slot;        // replica -> slot
    int32_t *slot_rep;        // slot -> replica
    const int32_t *lad_start; // nlad + 1 slot offsets
    double *E;                // per replica: Hamiltonian at the last measurement
    uint32_t *istep;          // next sweep index
    int32_t *row;             // next measurement row
    unsigned long long *swap_att, *swap_acc;  // per slot pair (t, t+1)
    double *part;             // per block: 8 measurement partials
    long long *spart;         // per block: spin sum
    double *rows;             // max_rows x nrep x kEnsCols
    int32_t max_rows, nrep, nlad, adapt;
    uint64_t seed;
    uint32_t run_index, K;
};

// KIND 0: director trials of `colour`, 1: chirality trials of `colour`; FP64
// only, 2: colour-1 director trials each followed by the site's chirality
// trial, 3: colour-0 chirality trials of sweep t each followed by the site's
// director trial of sweep t + 1 (the single-lattice fused launches).  One
// kernel per kind, so each gets the register budget of its own trial.
// A launch covers gridDim.x consecutive blocks of the table (one chunk of
// replicas).  rev = 0: block b takes entry base + b; rev = 1: entry base - b,
// base being the chunk's last entry, so that the launch walks them last to
// first and starts on the replicas the launch before it finished with (still
// in the memory-side cache).  The host forms base, so the kernel needs no
// grid-size load.
template <int KIND, int ARITH, int BC, int BS>
__global__ __launch_bounds__(BS) void k_phase_ens(EnsDev E, int colour, uint32_t istep_off, int rev, uint32_t base)
{
    constexpr int kind = KIND;
    const uint32_t code = E.blk_rep[rev ? base - blockIdx.x : base + blockIdx.x];
    const uint32_t r = code >> 24, lb = code & 0xFFFFFFu;
    const RepDesc &D = E.R[r];
    const int s = E.rep_slot[r];
    PhaseArgs A;
    A.g = D.g; A.KK = D.KK; A.kbt = E.kbt[s]; A.inv_kbt = E.inv_kbt[s]; A.inv_kbt_spin = A.inv_kbt; A.d = E.d + s;
    A.screen = kScreenScale;
    A.kbt_spin = A.kbt; A.flipbits = nullptr; A.held = nullptr;
    A.u_phi = nullptr; A.u_acc = nullptr; A.u_spin = nullptr;
    A.seed = E.seed; A.run_index = E.run_index; A.replica = D.replica; A.istep = *E.istep + istep_off;
    A.istep_spin = A.istep;
    const uint32_t m = lb * BS + threadIdx.x;
    const size_t gs = (size_t)kSlots * kSlotStride;
    bool acc = false, flip = false;
    if constexpr (ARITH == ARITH_FP64 && KIND >= 2) {
        if (kind == 2) {
            if (m < D.G.nsub) acc = dir_spin_site<RNG_PHILOX, BC, false>(D.G, D.F, 1, A, m, flip);
            block_count2<BS>(acc, flip, D.counts + kGroupAcc1 * gs, D.counts + flip_group(A.istep, 1) * gs, lb);
            return;
        }
        if (kind == 3) {
            A.istep = A.istep_spin + 1;
            if (m < D.G.nsub) acc = spin_dir_site<BC, false>(D.G, D.F, A, m, flip);
            block_count2<BS>(acc, flip, D.counts + kGroupAcc0 * gs, D.counts + flip_group(A.istep_spin, 0) * gs,
                             lb);
            return;
        }
    }
    if (m < D.G.nsub) {
        if (kind == 0) acc = dir_site<ARITH, RNG_PHILOX, BC>(D.G, D.F, colour, A, m);
        else acc = spin_site<ARITH, RNG_PHILOX, BC>(D.G, D.F, colour, A, m);
    }
    block_count<BS>(acc, D.counts + (kind == 0 ? colour : flip_group(A.istep, colour)) * gs, lb);
}

// FP64 kinds of k_phase_ens with two sites per thread and 16-byte pair loads
// and stores (the single lattice's k_pair), blocks from blk_rep2: the fused
// kinds 2 and 3, and kinds 0 and 1 for colour 0 (a block's first director and
// last chirality launch).  Needs every replica to satisfy pairs_ok.
template <int KIND, int BC, int BS>
__global__ __launch_bounds__(BS) void k_pair_ens(EnsDev E, uint32_t istep_off, int rev, uint32_t base)
{
    // the entry also holds the block's count of live threads, so the bounds
    // test needs no descriptor load and the descriptor loads issue together
    const uint2 ent = E.blk_rep2[rev ? base - blockIdx.x : base + blockIdx.x];
    const uint32_t r = ent.x >> 24, lb = ent.x & 0xFFFFFFu;
    const RepDesc &D = E.R[r];
    const int s = E.rep_slot[r];
    PhaseArgs A;
    A.g = D.g; A.KK = D.KK; A.kbt = E.kbt[s]; A.inv_kbt = E.inv_kbt[s]; A.inv_kbt_spin = A.inv_kbt; A.d = E.d + s;
    A.screen = kScreenScale;
    A.kbt_spin = A.kbt; A.flipbits = nullptr; A.held = nullptr;
    A.u_phi = nullptr; A.u_acc = nullptr; A.u_spin = nullptr;
    A.seed = E.seed; A.run_index = E.run_index; A.replica = D.replica; A.istep = *E.istep + istep_off;
    A.istep_spin = A.istep;
    const uint32_t m0 = 2u * (lb * BS + threadIdx.x);
    const size_t gs = (size_t)kSlots * kSlotStride;
    bool a0 = false, a1 = false, f0 = false, f1 = false;
    if constexpr (KIND == 3) A.istep = A.istep_spin + 1;
    const double d = KIND == 1 ? 0.0 : *A.d;  // the cone step, read once for both sites
    if (threadIdx.x < ent.y) {
        SiteIn S0, S1;
        const uint32_t own = load_pair<BC, true>(D.G, D.F, KIND == 2 ? 1 : 0, m0, S0, S1);
        SiteOut O0, O1;
        if constexpr (KIND == 0) {
            a0 = dir_apply<ARITH_FP64, RNG_PHILOX, true, true>(D.F, 0, A, d, m0, own, S0, &O0);
            a1 = dir_apply<ARITH_FP64, RNG_PHILOX, true, true>(D.F, 0, A, d, m0 + 1u, own + 1u, S1, &O1);
        } else if constexpr (KIND == 1) {
            f0 = spin_apply<ARITH_FP64, RNG_PHILOX, true, true>(D.F, 0, A, m0, own, S0, &O0);
            f1 = spin_apply<ARITH_FP64, RNG_PHILOX, true, true>(D.F, 0, A, m0 + 1u, own + 1u, S1, &O1);
        } else if constexpr (KIND == 2) {
            a0 = dir_spin_apply<RNG_PHILOX, true, true>(D.F, 1, A, d, m0, own, S0, f0, &O0);
            a1 = dir_spin_apply<RNG_PHILOX, true, true>(D.F, 1, A, d, m0 + 1u, own + 1u, S1, f1, &O1);
        } else {
            a0 = spin_dir_apply<true, true>(D.F, A, d, m0, own, S0, f0, &O0);
            a1 = spin_dir_apply<true, true>(D.F, A, d, m0 + 1u, own + 1u, S1, f1, &O1);
        }
        store_pair<KIND != 1, KIND != 0, true>(D.F, own, O0, O1);
    }
    const uint32_t wa = (uint32_t)(__popcll(__ballot(a0)) + __popcll(__ballot(a1)));
    const uint32_t wf = (uint32_t)(__popcll(__ballot(f0)) + __popcll(__ballot(f1)));
    if constexpr (KIND == 0)
        block_sum2<BS>(wa, 0u, D.counts + kGroupAcc0 * gs, D.counts + kGroupAcc0 * gs, lb);
    else if constexpr (KIND == 1)
        block_sum2<BS>(0u, wf, D.counts + flip_group(A.istep, 0) * gs, D.counts + flip_group(A.istep, 0) * gs, lb);
    else if constexpr (KIND == 2)
        block_sum2<BS>(wa, wf, D.counts + kGroupAcc1 * gs, D.counts + flip_group(A.istep, 1) * gs, lb);
    else
        block_sum2<BS>(wa, wf, D.counts + kGroupAcc0 * gs, D.counts + flip_group(A.istep_spin, 0) * gs, lb);
}

// Reference-arithmetic kinds of k_pair_ens, fused through held trials
// (chll_site.h): 0 director trials of colour 0 (a block's first launch), 2
// colour-1 director trials each followed by the site's screened chirality
// trial (held trials to the replica's queue, flips to its bitmap), 3 colour-0
// chirality trials of sweep t each followed by the site's director trial of
// sweep t + 1, 4 colour-0 chirality trials alone after colour 1's (a block's
// last sweep).  Needs every replica to satisfy pairs_ok.
// KIND 3 fits 3 waves per SIMD with a few spilled registers (the single
// lattice's k_pair<3> in reference arithmetic: 9 % faster than at 2).
template <int KIND, int BC, int BS>
__global__ __launch_bounds__(BS, (KIND == 3 ? 3 : 1)) void k_pair_ens_h(EnsDev E, uint32_t istep_off, int rev,
                                                                        uint32_t base)
{
    const uint2 ent = E.blk_rep2[rev ? base - blockIdx.x : base + blockIdx.x];
    const uint32_t r = ent.x >> 24, lb = ent.x & 0xFFFFFFu;
    const RepDesc &D = E.R[r];
    const int s = E.rep_slot[r];
    PhaseArgs A;
    A.g = D.g; A.KK = D.KK; A.kbt = E.kbt[s]; A.inv_kbt = E.inv_kbt[s]; A.inv_kbt_spin = A.inv_kbt; A.d = E.d + s;
    A.screen = kScreenScale;
    A.kbt_spin = A.kbt; A.flipbits = D.flipbits; A.held = D.held;
    A.u_phi = nullptr; A.u_acc = nullptr; A.u_spin = nullptr;
    A.seed = E.seed; A.run_index = E.run_index; A.replica = D.replica; A.istep = *E.istep + istep_off;
    A.istep_spin = A.istep;
    const uint32_t m0 = 2u * (lb * BS + threadIdx.x);
    const size_t gs = (size_t)kSlots * kSlotStride;
    bool a0 = false, a1 = false, f0 = false, f1 = false, h0 = false, h1 = false;
    if constexpr (KIND == 3) A.istep = A.istep_spin + 1;
    const double d = KIND == 4 ? 0.0 : *A.d;
    if (threadIdx.x < ent.y) {
        SiteIn S0, S1;
        const uint32_t own = load_pair<BC, true>(D.G, D.F, KIND == 2 ? 1 : 0, m0, S0, S1);
        SiteOut O0, O1;
        if constexpr (KIND == 0) {
            a0 = dir_apply<ARITH_COMPAT, RNG_PHILOX, true, true>(D.F, 0, A, d, m0, own, S0, &O0);
            a1 = dir_apply<ARITH_COMPAT, RNG_PHILOX, true, true>(D.F, 0, A, d, m0 + 1u, own + 1u, S1, &O1);
        } else if constexpr (KIND == 2) {
            a0 = dir_spin_held<RNG_PHILOX>(A, d, m0, S0, f0, h0, &O0);
            a1 = dir_spin_held<RNG_PHILOX>(A, d, m0 + 1u, S1, f1, h1, &O1);
        } else if constexpr (KIND == 3) {
            a0 = spin_dir_held<BC>(D.G, A, d, m0, S0, f0, &O0);
            a1 = spin_dir_held<BC>(D.G, A, d, m0 + 1u, S1, f1, &O1);
        } else {
            f0 = spin_c0_apply<RNG_PHILOX, BC>(D.G, A, m0, S0, &O0);
            f1 = spin_c0_apply<RNG_PHILOX, BC>(D.G, A, m0 + 1u, S1, &O1);
        }
        store_pair<KIND != 4, KIND != 0, true>(D.F, own, O0, O1);
    }
    if constexpr (KIND == 2) held_epilogue(A, D.G.nsub, true, lb * (BS / 64u) + (threadIdx.x >> 6), m0, f0, f1, h0, h1);
    const uint32_t wa = (uint32_t)(__popcll(__ballot(a0)) + __popcll(__ballot(a1)));
    const uint32_t wf = (uint32_t)(__popcll(__ballot(f0)) + __popcll(__ballot(f1)));
    if constexpr (KIND == 0)
        block_sum2<BS>(wa, 0u, D.counts + kGroupAcc0 * gs, D.counts + kGroupAcc0 * gs, lb);
    else if constexpr (KIND == 4)
        block_sum2<BS>(0u, wf, D.counts + flip_group(A.istep, 0) * gs, D.counts + flip_group(A.istep, 0) * gs, lb);
    else if constexpr (KIND == 2)
        block_sum2<BS>(wa, wf, D.counts + kGroupAcc1 * gs, D.counts + flip_group(A.istep, 1) * gs, lb);
    else
        block_sum2<BS>(wa, wf, D.counts + kGroupAcc0 * gs, D.counts + flip_group(A.istep_spin, 0) * gs, lb);
}

template <int BC, int BS>
__global__ __launch_bounds__(BS) void k_measure_ens(EnsDev E, uint32_t base)
{
    const uint32_t gb = base + blockIdx.x;
    const uint32_t code = E.blk_rep[gb];
    const uint32_t r = code >> 24, lb = code & 0xFFFFFFu;
    const RepDesc &D = E.R[r];
    MeasureAcc acc;
    const uint32_t m = lb * BS + threadIdx.x;
    if (m < D.G.nsub) measure_site<BC>(D.G, D.F, D.g, D.KK, m, acc);
    block_reduce_measure<BS>(acc, E.part + (size_t)gb * 8, E.spart + gb);
}

// The same blocks and partials with four sites of one row per thread
// (measure_pair): the first BS / 2 threads of a block take its BS packed
// indices as pairs.  Every replica has an even half and fields below 4 GiB.
template <int BC, int BS>
__global__ __launch_bounds__(BS) void k_measure_ens_pair(EnsDev E, uint32_t base)
{
    const uint32_t gb = base + blockIdx.x;
    const uint32_t code = E.blk_rep[gb];
    const uint32_t r = code >> 24, lb = code & 0xFFFFFFu;
    const RepDesc &D = E.R[r];
    MeasureAcc acc;
    const uint32_t beg = lb * BS;
    const uint32_t end = beg + BS < D.G.nsub ? beg + BS : D.G.nsub;
    const uint32_t m0 = beg + 2u * threadIdx.x;
    if (threadIdx.x < BS / 2 && m0 < end)
        measure_pair<BC>(D.G, D.F, D.g, D.KK, m0, (threadIdx.x & 63u) != 63u && m0 + 2u < end, acc);
    block_reduce_measure<BS>(acc, E.part + (size_t)gb * 8, E.spart + gb);
}

// The same observables marching along k (measure_march_column, as the single
// lattice's k_measure_march): block lb of replica r takes tile lb / nslab of
// BS packed pairs of a plane through slab lb % nslab of 8 planes, so the
// blocks in flight cover all slabs of a few tiles; its partial is the
// replica's lb-th.  Used when every replica's march blocks fit its
// partials (nmeas <= nblk).
template <int BC, int BS>
__global__ __launch_bounds__(BS) void k_measure_ens_march(EnsDev E, uint32_t base)
{
    const uint32_t code = E.blk_rep3[base + blockIdx.x];
    const uint32_t r = code >> 24, lb = code & 0xFFFFFFu;
    const RepDesc &D = E.R[r];
    const uint32_t quads = D.G.plane >> 1;
    const uint32_t tiles = (quads + BS - 1) / BS;
    const uint32_t ks = D.G.nk < 8u ? D.G.nk : 8u;
    const uint32_t nslab = (D.G.nk + ks - 1) / ks;
    const uint32_t t = lb / nslab, sl = lb - t * nslab;  // tile by tile through the slabs
    const uint32_t qd = t * BS + threadIdx.x;
    const uint32_t k0 = sl * ks, k1 = k0 + ks < D.G.nk ? k0 + ks : D.G.nk;
    MeasureAcc acc;
    if (qd < quads && k0 < k1) {
        const bool nxt = (threadIdx.x & 63u) != 63u && qd + 1u < quads;
        measure_march_column<BC>(D.G, D.F, D.g, D.KK, qd, k0, k1, (sl & 1u) == 0u, nxt, acc);
    }
    const size_t gb = (size_t)D.blk_start + lb;
    block_reduce_measure<BS>(acc, E.part + gb * 8, E.spart + gb);
}

// Per-replica epilogue (one block per replica): accept/flip totals, step
// adaptation of the replica's temperature slot (chll/chll_mod.f90:83-90), and
// on measuring sweeps the observable row and the replica's energy for the swap.
constexpr int kEnsFin = 256;

// The director accepts of sweep istep are always closed (cone-step update);
// close_prev zeroes the flips of sweep istep - 1, close_this counts those of
// sweep istep (for the row) and zeroes them.
// The per-replica epilogue of replica order[fbase + b], one block of kEnsFin
// threads (k_finalize_ens, and blocks of k_held_ens).
__device__ inline void finalize_ens_block(const EnsDev &E, uint32_t istep_off, int measure, int close_prev,
                                          int close_this, uint32_t fbase, uint32_t b)
{
    const int r = (int)E.order[fbase + b];
    const RepDesc &D = E.R[r];
    const uint32_t t = threadIdx.x;
    const uint32_t istep = *E.istep + istep_off;
    const int gt = flip_group((int32_t)istep, 0), gp = flip_group((int32_t)istep - 1, 0);
    unsigned long long ca = 0, cf = 0;
    // every slot a thread owns is read before any is cleared, so its loads
    // are in flight together
    constexpr int kPer = (kGroups * kSlots + kEnsFin - 1) / kEnsFin;
    uint32_t *slot[kPer];
    uint32_t cv[kPer];
    int use[kPer];  // 0 none, 1 accepts, 2 this sweep's flips, 3 the previous sweep's (cleared only)
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const uint32_t q = t + (uint32_t)i * kEnsFin;
        const int g = (int)(q / kSlots);
        use[i] = q >= (uint32_t)(kGroups * kSlots) ? 0
                 : g < 2                                                ? 1
                 : close_this && (g == gt || g == gt + 1)               ? 2
                 : close_prev && (g == gp || g == gp + 1)               ? 3
                                                                        : 0;
        slot[i] = use[i] ? D.counts + g * (kSlots * kSlotStride) + (q % kSlots) * kSlotStride : D.counts;
        cv[i] = use[i] ? *slot[i] : 0u;
    }
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        if (!use[i]) continue;
        *slot[i] = 0u;
        if (use[i] == 1) ca += cv[i];
        else if (use[i] == 2) cf += cv[i];
    }
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = 0.0;
    long long ss = 0;
    if (measure) {
        for (uint32_t b = t; b < D.nmeas; b += kEnsFin) {
            const size_t gb = (size_t)D.blk_start + b;
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] += E.part[gb * 8 + q];
            ss += E.spart[gb];
        }
    }
    double cad = (double)ca, cfd = (double)cf, ssd = (double)ss;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        cad += __shfl_xor(cad, off, 64);
        cfd += __shfl_xor(cfd, off, 64);
    }
    if (measure) {  // uniform: sweeps without a row skip the observable shuffles
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] += __shfl_xor(v[q], off, 64);
            ssd += __shfl_xor(ssd, off, 64);
        }
    }
    __shared__ double red[kEnsFin / 64][11];
    const uint32_t lane = t & 63u, wid = t >> 6;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) red[wid][q] = v[q];
        red[wid][8] = cad;
        red[wid][9] = cfd;
        red[wid][10] = ssd;
    }
    __syncthreads();
    if (t != 0) return;
    double tot[11];
#pragma unroll
    for (int q = 0; q < 11; ++q) {
        double a = 0.0;
#pragma unroll
        for (int w = 0; w < kEnsFin / 64; ++w) a += red[w][q];
        tot[q] = a;
    }
    const int s = E.rep_slot[r];
    const double n3 = 2.0 * (double)D.G.nsub;
    const double n3f = (double)(float)n3;
    const double paccept = (double)(float)((double)(float)(int32_t)(long long)tot[8] / n3f);
    double d = E.d[s];
    if (E.adapt && istep < 10000) {
        if (paccept < 0.4) d = d * 0.99500000476837158203125;
        else if (paccept > 0.6) d = d / 0.99500000476837158203125;
    }
    E.d[s] = d;
    if (measure) {
        E.E[r] = tot[1];
        const int row = *E.row;
        if (row < E.max_rows) {
            double *o = E.rows + ((size_t)row * E.nrep + r) * kEnsCols;
            o[0] = (double)istep;
            o[1] = E.kbt[s];
            o[2] = tot[1] / n3;
            o[3] = tot[10] / n3;
            o[4] = tot[2] / n3;
            o[5] = tot[3] / n3;
            o[6] = tot[4] / n3;
            o[7] = tot[5] / n3;
            o[8] = tot[6] / n3;
            o[9] = tot[7] / n3;
            o[10] = tot[8] / n3;
            o[11] = tot[9] / n3;
            o[12] = d;
            o[13] = (double)s;
            o[14] = (double)(float)(tot[0] / n3f);
            o[15] = 0.0;
        }
    }
}

__global__ __launch_bounds__(kEnsFin) void k_finalize_ens(EnsDev E, uint32_t istep_off, int measure, int close_prev,
                                                          int close_this, uint32_t fbase)
{
    finalize_ens_block(E, istep_off, measure, close_prev, close_this, fbase, blockIdx.x);
}

// Held chirality trials (held_group) of the nf replicas order[first ..
// first + nf - 1], hb blocks per replica.  With fin, nf more blocks run those
// replicas' epilogue that closes the sweep's director accepts
// (finalize_ens_block, close_prev as given): it reads the accepts and the
// previous sweep's flips, none of which the held trials write.
constexpr uint32_t kHeldEnsBlocks = 2;
static_assert(kHeldBlockThreads == kEnsFin, "the epilogue blocks of k_held_ens take the held blocks' size");
template <int BC>
__global__ __launch_bounds__(kHeldBlockThreads) void k_held_ens(EnsDev E, uint32_t first, uint32_t nf, uint32_t hb,
                                                              uint32_t istep_off, int fin, int close_prev)
{
    if (blockIdx.x >= nf * hb) {
        if (fin) finalize_ens_block(E, istep_off, 0, close_prev, 0, first, blockIdx.x - nf * hb);
        return;
    }
    const uint32_t r = E.order[first + blockIdx.x / hb], blk = blockIdx.x % hb;
    const RepDesc &D = E.R[r];
    const int s = E.rep_slot[r];
    PhaseArgs A;
    A.g = D.g; A.KK = D.KK; A.kbt = E.kbt[s]; A.inv_kbt = E.inv_kbt[s]; A.inv_kbt_spin = A.inv_kbt; A.d = E.d + s;
    A.screen = kScreenScale;
    A.kbt_spin = A.kbt; A.flipbits = D.flipbits; A.held = D.held;
    A.u_phi = nullptr; A.u_acc = nullptr; A.u_spin = nullptr;
    A.seed = E.seed; A.run_index = E.run_index; A.replica = D.replica; A.istep = *E.istep + istep_off;
    A.istep_spin = A.istep;
    const size_t gs = (size_t)kSlots * kSlotStride;
    held_group<RNG_PHILOX, BC>(D.G, D.F, A, A, D.held, D.counts + flip_group(A.istep, 1) * gs, nullptr, blk, hb);
}


// Parallel-tempering swaps, one thread per ladder, then advance the sweep and
// row counters.  Round q = istep / K alternates even and odd slot pairs.
__global__ void k_swap_ens(EnsDev E, int do_swap)
{
    const int l = threadIdx.x;
    const uint32_t istep_last = *E.istep + E.K - 1;
    __syncthreads();
    if (do_swap && l < E.nlad) {
        const uint32_t round = istep_last / E.K;
        const int t0 = E.lad_start[l], t1 = E.lad_start[l + 1];
        for (int t = t0 + (int)(round & 1u); t + 1 < t1; t += 2) {
            const int a = E.slot_rep[t], b = E.slot_rep[t + 1];
            const double delta = (E.inv_kbt[t] - E.inv_kbt[t + 1]) * (E.E[a] - E.E[b]);
            const U4 w = philox4x32_10(U4{(uint32_t)t, round, E.run_index, 0x50540000u}, (uint32_t)E.seed,
                                       (uint32_t)(E.seed >> 32));
            const double u = u53(w.y, w.x);
            E.swap_att[t] += 1ull;
            if (delta >= 0.0 || u < exp(delta)) {
                E.slot_rep[t] = b;
                E.slot_rep[t + 1] = a;
                E.rep_slot[a] = t + 1;
                E.rep_slot[b] = t;
                E.swap_acc[t] += 1ull;
            }
        }
    }
    __syncthreads();
    if (l == 0) {
        *E.istep = istep_last + 1;
        *E.row = *E.row + 1;
    }
}

}  // namespace chll

using namespace chll;

namespace {

thread_local std::string e_err;

int efail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    e_err = buf;
    return -1;
}

#define ECHECK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) return efail("%s: %s", #x, hipGetErrorString(e_));       \
    } while (0)

}  // namespace

// One chunk of replicas (consecutive in the block-table order): its blocks in
// the one-site table [t1, t1 + n1), in the pair table [t2, t2 + n2) and in the
// marching-observables table [t3, t3 + n3), its replicas order[f .. f + nf).
struct Chunk {
    uint32_t t1, n1, t2, n2, t3, n3, f, nf;
};

struct chll_ensemble {
    int nrep = 0, nlad = 0, bc = BC_OPEN, arith = ARITH_COMPAT, bs = 256;
    std::vector<chll_replica_desc> rd;
    std::vector<RepDesc> rep;       // host copy of descriptors
    std::vector<int32_t> lad_start;
    std::vector<uint32_t> order;    // replica order of blocks (largest first)
    hipStream_t own = nullptr, stream = nullptr;
    RepDesc *R = nullptr;
    uint32_t *blk_rep = nullptr;
    uint32_t nblocks = 0;
    uint2 *blk_rep2 = nullptr;  // pair-kernel blocks (when every replica allows site pairs)
    uint32_t nblocks2 = 0;
    bool pairs = false;
    uint32_t *blk_rep3 = nullptr;  // marching-observables blocks (pairs, and every replica's fit its partials)
    uint32_t nblocks3 = 0;
    bool march = false;
    uint32_t *order_dev = nullptr;
    // Replicas are independent between swaps, so a block of K sweeps runs
    // chunk by chunk: all K sweeps of one chunk of at most chunk_bytes of state
    // (0: one chunk), which then stays in the memory-side cache.
    size_t chunk_bytes = 0;
    std::vector<Chunk> chunks;
    // Chunks are independent, so with branches = n > 1 they take turns on n
    // streams forked from and joined back to the run's stream (n branches of
    // the captured graph): one chunk's short launches (epilogues, tails) run
    // beside another's sweeps.
    static constexpr int kMaxBranches = 3;
    int branches = 1;
    hipStream_t br[kMaxBranches] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kMaxBranches] = {};
    EnsDev E{};
    double *stage = nullptr;
    int32_t *stage_s = nullptr;
    int *bad = nullptr;
    size_t stage_n = 0;
    int32_t rows_cap = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_K = 0, graph_bs = 0, graph_adapt = -1, graph_fuse = -1;
    int fuse = 1;  // FP64: 0 four half-sweeps, 1 fused launches (site pairs where allowed), 2 fused, one site
    uint64_t graph_seed = 0;
    uint32_t graph_run = 0;
    double *pinned = nullptr;
};

namespace {

template <class T>
int ealloc(T **p, size_t n)
{
    *p = nullptr;
    hipError_t e = hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess) return efail("hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
    return 0;
}

void drop_graph(chll_ensemble *X)
{
    if (X->gexec) (void)hipGraphExecDestroy(X->gexec);
    if (X->graph) (void)hipGraphDestroy(X->graph);
    X->gexec = nullptr;
    X->graph = nullptr;
    X->graph_K = 0;
}

// Chunks of consecutive replicas in block-table order, each holding at most
// X->chunk_bytes of lattice state (a replica larger than that is a chunk alone).
void build_chunks(chll_ensemble *X)
{
    X->chunks.clear();
    Chunk c{0, 0, 0, 0, 0, 0, 0, 0};
    size_t bytes = 0;
    uint32_t p1 = 0, p2 = 0, p3 = 0;
    for (uint32_t i = 0; i < (uint32_t)X->order.size(); ++i) {
        const RepDesc &D = X->rep[X->order[i]];
        const size_t rb = (2 * (size_t)D.G.nsub + kFieldPad) * (3 * sizeof(double) + 1);
        const uint32_t nb1 = D.nblk;
        const uint32_t nb2 = X->pairs ? (D.G.nsub / 2 + X->bs - 1) / X->bs : 0u;
        const uint32_t nb3 = X->march ? D.nmeas : 0u;
        if (c.nf > 0 && X->chunk_bytes > 0 && bytes + rb > X->chunk_bytes) {
            X->chunks.push_back(c);
            c = Chunk{p1, 0, p2, 0, p3, 0, i, 0};
            bytes = 0;
        }
        c.n1 += nb1;
        c.n2 += nb2;
        c.n3 += nb3;
        c.nf += 1;
        bytes += rb;
        p1 += nb1;
        p2 += nb2;
        p3 += nb3;
    }
    if (c.nf) X->chunks.push_back(c);
}

// (Re)build the block table and descriptors for block size X->bs.
int build_tables(chll_ensemble *X)
{
    std::vector<uint32_t> table;
    uint32_t start = 0;
    for (uint32_t r : X->order) {
        RepDesc &D = X->rep[r];
        D.nblk = (D.G.nsub + X->bs - 1) / X->bs;
        D.blk_start = start;
        if (D.nblk >= (1u << 24)) return efail("replica %u too large for the block table", r);
        for (uint32_t b = 0; b < D.nblk; ++b) table.push_back((r << 24) | b);
        start += D.nblk;
    }
    X->nblocks = start;
    // Pair-kernel blocks.  A replica whose pair-plane splits into 8 bands of
    // whole blocks, starting at a table position that is a multiple of 8, gets
    // XCD row bands: table position p runs on XCD p % 8 (reversed launches
    // permute the XCDs the same way for every block), and the band_block order
    // gives that XCD one band of rows of every plane, so the y neighbours of a
    // row stay in one L2.  Only the integer accept / flip counts depend on
    // these blocks.
    std::vector<uint2> table2;
    if (X->pairs) {
        for (uint32_t r : X->order) {
            const Geom &G = X->rep[r].G;
            const uint32_t nb = (G.nsub / 2 + X->bs - 1) / X->bs;
            const uint32_t per = 8u * (uint32_t)X->bs;
            const uint32_t band =
                (G.plane / 2) % per == 0 && table2.size() % 8 == 0 ? (G.plane / 2) / per : 0u;
            for (uint32_t b = 0; b < nb; ++b) {
                uint32_t lb = b;
                if (band) {
                    const uint32_t x = b & 7u, n = b >> 3, q = n / band, rr = n - q * band;
                    lb = (q * 8u + x) * band + rr;
                }
                const uint32_t live = std::min<uint32_t>((uint32_t)X->bs, G.nsub / 2 - lb * (uint32_t)X->bs);
                table2.push_back(make_uint2((r << 24) | lb, live));
            }
        }
    }
    X->nblocks2 = (uint32_t)table2.size();
    // Marching-observables blocks: per replica tiles of bs packed pairs of a
    // plane x slabs of 8 planes, wherever every replica's count fits its
    // one-site blocks (whose partial slots the measure writes).
    std::vector<uint32_t> table3;
    X->march = true;  // in either arithmetic: the observables are binary64 in both
    for (uint32_t r : X->order) {
        const Geom &G = X->rep[r].G;
        if (!pairs_ok(G)) X->march = false;
        const uint32_t tiles = ((G.plane >> 1) + X->bs - 1) / X->bs;
        const uint32_t ks = G.nk < 8u ? G.nk : 8u;
        const uint64_t nm = (uint64_t)tiles * ((G.nk + ks - 1) / ks);
        if (nm > X->rep[r].nblk) X->march = false;
    }
    for (uint32_t r : X->order) {
        RepDesc &D = X->rep[r];
        D.nmeas = D.nblk;
        if (!X->march) continue;
        const uint32_t tiles = ((D.G.plane >> 1) + X->bs - 1) / X->bs;
        const uint32_t ks = D.G.nk < 8u ? D.G.nk : 8u;
        D.nmeas = tiles * ((D.G.nk + ks - 1) / ks);
        for (uint32_t b = 0; b < D.nmeas; ++b) table3.push_back((r << 24) | b);
    }
    X->nblocks3 = (uint32_t)table3.size();
    if (X->blk_rep) (void)hipFree(X->blk_rep);
    if (X->blk_rep2) (void)hipFree(X->blk_rep2);
    if (X->blk_rep3) (void)hipFree(X->blk_rep3);
    X->blk_rep = nullptr;
    X->blk_rep2 = nullptr;
    X->blk_rep3 = nullptr;
    if (X->E.part) (void)hipFree(X->E.part);
    if (X->E.spart) (void)hipFree(X->E.spart);
    X->E.part = nullptr;
    X->E.spart = nullptr;
    if (ealloc(&X->blk_rep, table.size()) || ealloc(&X->E.part, (size_t)X->nblocks * 8) ||
        ealloc(&X->E.spart, X->nblocks) || (X->pairs && ealloc(&X->blk_rep2, table2.size())) ||
        (X->march && ealloc(&X->blk_rep3, table3.size())))
        return -1;
    ECHECK(hipMemcpy(X->blk_rep, table.data(), table.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (X->pairs)
        ECHECK(hipMemcpy(X->blk_rep2, table2.data(), table2.size() * sizeof(uint2), hipMemcpyHostToDevice));
    if (X->march)
        ECHECK(hipMemcpy(X->blk_rep3, table3.data(), table3.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    X->E.blk_rep2 = X->blk_rep2;
    X->E.blk_rep3 = X->blk_rep3;
    ECHECK(hipMemcpy(X->R, X->rep.data(), X->rep.size() * sizeof(RepDesc), hipMemcpyHostToDevice));
    if (!X->order_dev && ealloc(&X->order_dev, X->order.size())) return -1;
    ECHECK(hipMemcpy(X->order_dev, X->order.data(), X->order.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    X->E.R = X->R;
    X->E.blk_rep = X->blk_rep;
    X->E.order = X->order_dev;
    build_chunks(X);
    drop_graph(X);
    return 0;
}

// One block of K sweeps.  FP64 with fusion on, as the single lattice: D0 of
// the first sweep alone, then per sweep t the fused D1 + C1 launch and either
// (t < K - 1) the epilogue closing t's director accepts and the fused C0(t) +
// D0(t + 1) launch, or (last sweep, which is measured) C0 alone, the
// observable reduction and the closing epilogue.  The swap kernel then moves
// replicas between temperature slots, so no fused launch crosses it.
template <int ARITH, int BC, int BS>
int launch_block_t(chll_ensemble *X, int K, hipStream_t st)
{
    const EnsDev &E = X->E;
    const int nbr = std::min<int>(X->branches, (int)X->chunks.size());
    hipStream_t cst[chll_ensemble::kMaxBranches] = {st};
    if (nbr > 1) {
        ECHECK(hipEventRecord(X->ev_fork, st));
        for (int q = 0; q < nbr; ++q) {
            ECHECK(hipStreamWaitEvent(X->br[q], X->ev_fork, 0));
            cst[q] = X->br[q];
        }
    }
    const dim3 b(BS), bf(kEnsFin);
    // reference arithmetic fuses only through held trials, on site pairs
    const bool hfuse = ARITH == ARITH_COMPAT && X->fuse == 1 && X->pairs;
    const bool fuse = (ARITH == ARITH_FP64 && X->fuse != 0) || hfuse;
    const bool pairs = fuse && X->fuse == 1 && X->pairs;
    // observables: four sites per thread wherever every replica allows site
    // pairs, whatever the fusion setting, so rows do not depend on it
    auto measure = [&](const Chunk &c, hipStream_t ms) {
        if (X->march) hipLaunchKernelGGL((k_measure_ens_march<BC, BS>), dim3(c.n3), dim3(BS), 0, ms, E, c.t3);
        else if (X->pairs) hipLaunchKernelGGL((k_measure_ens_pair<BC, BS>), dim3(c.n1), dim3(BS), 0, ms, E, c.t1);
        else hipLaunchKernelGGL((k_measure_ens<BC, BS>), dim3(c.n1), dim3(BS), 0, ms, E, c.t1);
    };
    for (size_t ci = 0; ci < X->chunks.size(); ++ci) {
        const Chunk &c = X->chunks[ci];
        const hipStream_t cs = cst[ci % nbr];  // this chunk's branch
        const dim3 g(c.n1), g2(c.n2), gf(c.nf);
        const uint32_t r1 = c.t1 + c.n1 - 1u, r2 = c.t2 + c.n2 - 1u;  // last entries, for reversed walks
        for (int t = 0; t < K; ++t) {
            const uint32_t off = (uint32_t)t;
            const bool last = t + 1 == K;
            // consecutive sweep launches alternate their block order
            if (!fuse) {
                hipLaunchKernelGGL((k_phase_ens<0, ARITH, BC, BS>), g, b, 0, cs, E, 0, off, 0, c.t1);
                hipLaunchKernelGGL((k_phase_ens<0, ARITH, BC, BS>), g, b, 0, cs, E, 1, off, 1, r1);
                hipLaunchKernelGGL((k_phase_ens<1, ARITH, BC, BS>), g, b, 0, cs, E, 0, off, 0, c.t1);
                hipLaunchKernelGGL((k_phase_ens<1, ARITH, BC, BS>), g, b, 0, cs, E, 1, off, 1, r1);
                if (last) measure(c, cs);
                hipLaunchKernelGGL(k_finalize_ens, gf, bf, 0, cs, E, off, last ? 1 : 0, 0, 1, c.f);
                continue;
            }
            if (t == 0) {
                if (hfuse) hipLaunchKernelGGL((k_pair_ens_h<0, BC, BS>), g2, b, 0, cs, E, off, 1, r2);
                else if (pairs) hipLaunchKernelGGL((k_pair_ens<0, BC, BS>), g2, b, 0, cs, E, off, 1, r2);
                else hipLaunchKernelGGL((k_phase_ens<0, ARITH, BC, BS>), g, b, 0, cs, E, 0, off, 1, r1);
            }
            if (hfuse) {
                // colour 1's director + chirality trials, then its held trials
                hipLaunchKernelGGL((k_pair_ens_h<2, BC, BS>), g2, b, 0, cs, E, off, 0, c.t2);
                // with, in a chain, the epilogue closing the director accepts
                hipLaunchKernelGGL((k_held_ens<BC>), dim3(c.nf * (kHeldEnsBlocks + (last ? 0u : 1u))),
                                   dim3(kHeldBlockThreads), 0, cs, E, c.f, c.nf, kHeldEnsBlocks, off, last ? 0 : 1,
                                   t > 0 ? 1 : 0);
            } else if (pairs) {
                hipLaunchKernelGGL((k_pair_ens<2, BC, BS>), g2, b, 0, cs, E, off, 0, c.t2);
            } else {
                hipLaunchKernelGGL((k_phase_ens<2, ARITH, BC, BS>), g, b, 0, cs, E, 1, off, 0, c.t1);
            }
            if (!last) {
                if (!hfuse) hipLaunchKernelGGL(k_finalize_ens, gf, bf, 0, cs, E, off, 0, t > 0 ? 1 : 0, 0, c.f);
                if (hfuse) hipLaunchKernelGGL((k_pair_ens_h<3, BC, BS>), g2, b, 0, cs, E, off, 1, r2);
                else if (pairs) hipLaunchKernelGGL((k_pair_ens<3, BC, BS>), g2, b, 0, cs, E, off, 1, r2);
                else hipLaunchKernelGGL((k_phase_ens<3, ARITH, BC, BS>), g, b, 0, cs, E, 0, off, 1, r1);
            } else {
                if (hfuse) hipLaunchKernelGGL((k_pair_ens_h<4, BC, BS>), g2, b, 0, cs, E, off, 1, r2);
                else if (pairs) hipLaunchKernelGGL((k_pair_ens<1, BC, BS>), g2, b, 0, cs, E, off, 1, r2);
                else hipLaunchKernelGGL((k_phase_ens<1, ARITH, BC, BS>), g, b, 0, cs, E, 0, off, 1, r1);
                measure(c, cs);
                hipLaunchKernelGGL(k_finalize_ens, gf, bf, 0, cs, E, off, 1, t > 0 ? 1 : 0, 1, c.f);
            }
        }
    }
    if (nbr > 1) {
        for (int q = 0; q < nbr; ++q) {
            ECHECK(hipEventRecord(X->ev_join[q], X->br[q]));
            ECHECK(hipStreamWaitEvent(st, X->ev_join[q], 0));
        }
    }
    hipLaunchKernelGGL(k_swap_ens, dim3(1), dim3(256), 0, st, E, 1);
    return 0;
}

template <int ARITH, int BC>
int launch_block_bs(chll_ensemble *X, int K, hipStream_t st)
{
    switch (X->bs) {
    case 64: return launch_block_t<ARITH, BC, 64>(X, K, st);
    case 128: return launch_block_t<ARITH, BC, 128>(X, K, st);
    case 256: return launch_block_t<ARITH, BC, 256>(X, K, st);
    case 512: return launch_block_t<ARITH, BC, 512>(X, K, st);
    default: return efail("unsupported block size %d", X->bs);
    }
}

// Enqueue one block of K sweeps (+ measurement + swap).
int launch_block(chll_ensemble *X, int K, hipStream_t st)
{
    int rc;
    if (X->arith == ARITH_COMPAT) rc = X->bc == BC_OPEN ? launch_block_bs<ARITH_COMPAT, BC_OPEN>(X, K, st)
                                                        : launch_block_bs<ARITH_COMPAT, BC_PERIODIC>(X, K, st);
    else rc = X->bc == BC_OPEN ? launch_block_bs<ARITH_FP64, BC_OPEN>(X, K, st)
                               : launch_block_bs<ARITH_FP64, BC_PERIODIC>(X, K, st);
    if (rc) return rc;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return efail("ensemble launch: %s", hipGetErrorString(e));
    return 0;
}

}  // namespace

extern "C" {

const char *chll_ensemble_last_error(void) { return e_err.c_str(); }

int chll_ensemble_create(const chll_replica_desc *reps, int32_t nrep, int32_t bc, int32_t arith, chll_ensemble **out)
{
    e_err.clear();
    if (!reps || !out || nrep < 1 || nrep > 255) return efail("need 1..255 replicas");
    if (bc != BC_OPEN && bc != BC_PERIODIC) return efail("unknown boundary condition %d", bc);
    if (arith != ARITH_COMPAT && arith != ARITH_FP64) return efail("unknown arithmetic %d", arith);
    *out = nullptr;
    // validate ladders: contiguous, same lattice and Hamiltonian, ascending kT
    std::vector<int32_t> lad_start;
    for (int r = 0; r < nrep; ++r) {
        const chll_replica_desc &a = reps[r];
        if (a.ni < 2 || a.ni % 2 || a.nj < 1 || a.nk < 1)
            return efail("replica %d: dims must have ni even >= 2 (got %d,%d,%d)", r, a.ni, a.nj, a.nk);
        if (bc == BC_PERIODIC && ((a.nj % 2) || (a.nk % 2)))
            return efail("replica %d: periodic boundaries need even nj, nk", r);
        if ((uint64_t)a.ni * a.nj * a.nk + kFieldPad >= (1ull << 32)) return efail("replica %d too large", r);
        if (!(a.kbt > 0.0)) return efail("replica %d: kbt must be positive", r);
        if (r == 0 || a.ladder != reps[r - 1].ladder) {
            for (int q = 0; q < r; ++q)
                if (reps[q].ladder == a.ladder) return efail("replica %d: ladder %d is not contiguous", r, a.ladder);
            lad_start.push_back(r);
        } else {
            const chll_replica_desc &p = reps[r - 1];
            if (p.ni != a.ni || p.nj != a.nj || p.nk != a.nk || p.g != a.g || p.KK != a.KK)
                return efail("replica %d: a ladder needs one lattice size, g and K", r);
            if (!(a.kbt > p.kbt)) return efail("replica %d: kbt must ascend within a ladder", r);
        }
    }
    lad_start.push_back(nrep);
    chll_ensemble *X = new chll_ensemble();
    X->nrep = nrep;
    X->nlad = (int)lad_start.size() - 1;
    X->bc = bc;
    X->arith = arith;
    X->rd.assign(reps, reps + nrep);
    X->lad_start = lad_start;
    X->rep.resize(nrep);
    int rc = 0;
    for (int r = 0; r < nrep && !rc; ++r) {
        RepDesc &D = X->rep[r];
        memset(&D, 0, sizeof D);
        Geom &G = D.G;
        G.ni = reps[r].ni; G.nj = reps[r].nj; G.nk = reps[r].nk;
        G.half = G.ni / 2;
        G.plane = G.half * G.nj;
        G.nsub = (uint32_t)(((uint64_t)G.ni * G.nj * G.nk) / 2);
        G.zero = 2 * G.nsub;
        G.nplane = G.ni * G.nj;
        G.fd_half = make_fastdiv(G.half);
        G.fd_nj = make_fastdiv(G.nj);
        G.fd_ni = make_fastdiv(G.ni);
        const size_t n = 2 * (size_t)G.nsub + kFieldPad;
        rc |= ealloc(&D.F.nx, n) | ealloc(&D.F.ny, n) | ealloc(&D.F.nz, n) | ealloc(&D.F.s, n);
        rc |= ealloc(&D.counts, (size_t)kCountWords);
        if (!rc) {
            rc |= hipMemset(D.F.nx, 0, n * sizeof(double)) != hipSuccess;
            rc |= hipMemset(D.F.ny, 0, n * sizeof(double)) != hipSuccess;
            rc |= hipMemset(D.F.nz, 0, n * sizeof(double)) != hipSuccess;
            rc |= hipMemset(D.F.s, 0, n) != hipSuccess;
            rc |= hipMemset(D.counts, 0, kCountWords * sizeof(uint32_t)) != hipSuccess;
            if (rc) efail("hipMemset failed");
        }
        D.replica = (uint32_t)r;
        D.g = reps[r].g;
        D.KK = reps[r].KK;
    }
    // two sites per thread in the fused launches when every replica allows it
    X->pairs = true;
    for (int r = 0; r < nrep; ++r) X->pairs = X->pairs && pairs_ok(X->rep[r].G);
    // reference arithmetic fuses through held trials: a flip bitmap (two words
    // per wave of every block size) and a held-trial queue per replica
    for (int r = 0; r < nrep && !rc && arith == ARITH_COMPAT && X->pairs; ++r) {
        RepDesc &D = X->rep[r];
        const size_t words = 16 * (((size_t)D.G.nsub / 2 + 511) / 512);
        rc |= ealloc(&D.flipbits, words) | ealloc(&D.held, kHeldHead + (size_t)D.G.nsub);
        if (!rc && hipMemset(D.held, 0, kHeldHead * sizeof(uint32_t)) != hipSuccess) rc = efail("hipMemset failed");
    }
    // block order: largest replicas first
    X->order.resize(nrep);
    for (int r = 0; r < nrep; ++r) X->order[r] = r;
    std::stable_sort(X->order.begin(), X->order.end(),
                     [&](uint32_t a, uint32_t b) { return X->rep[a].G.nsub > X->rep[b].G.nsub; });
    EnsDev &E = X->E;
    E.nrep = nrep;
    E.nlad = X->nlad;
    rc = rc || ealloc(&X->R, nrep) || ealloc(&E.kbt, nrep) || ealloc(&E.inv_kbt, nrep) || ealloc(&E.d, nrep) ||
         ealloc(&E.rep_slot, nrep) || ealloc(&E.slot_rep, nrep) || ealloc((int32_t **)&E.lad_start, X->nlad + 1) ||
         ealloc(&E.E, nrep) || ealloc(&E.istep, 1) || ealloc(&E.row, 1) || ealloc(&E.swap_att, nrep) ||
         ealloc(&E.swap_acc, nrep) || ealloc(&X->bad, 1);
    if (!rc && hipHostMalloc((void **)&X->pinned, 16 * sizeof(double)) != hipSuccess) rc = efail("hipHostMalloc");
    if (!rc && hipStreamCreate(&X->own) != hipSuccess) rc = efail("hipStreamCreate");
    if (!rc) {
        std::vector<double> kb(nrep), ikb(nrep), d(nrep);
        std::vector<int32_t> ident(nrep);
        for (int r = 0; r < nrep; ++r) {
            kb[r] = reps[r].kbt;
            ikb[r] = 1.0 / reps[r].kbt;
            d[r] = reps[r].d0;
            ident[r] = r;
        }
        uint32_t one = 1;
        int32_t zero = 0;
        std::vector<unsigned long long> z(nrep, 0ull);
        rc |= hipMemcpy(E.kbt, kb.data(), nrep * sizeof(double), hipMemcpyHostToDevice) != hipSuccess;
        rc |= hipMemcpy(E.inv_kbt, ikb.data(), nrep * sizeof(double), hipMemcpyHostToDevice) != hipSuccess;
        rc |= hipMemcpy(E.d, d.data(), nrep * sizeof(double), hipMemcpyHostToDevice) != hipSuccess;
        rc |= hipMemcpy(E.rep_slot, ident.data(), nrep * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess;
        rc |= hipMemcpy(E.slot_rep, ident.data(), nrep * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess;
        rc |= hipMemcpy((void *)E.lad_start, lad_start.data(), lad_start.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice) != hipSuccess;
        rc |= hipMemcpy(E.istep, &one, sizeof one, hipMemcpyHostToDevice) != hipSuccess;
        rc |= hipMemcpy(E.row, &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess;
        rc |= hipMemcpy(E.swap_att, z.data(), nrep * sizeof(unsigned long long), hipMemcpyHostToDevice) != hipSuccess;
        rc |= hipMemcpy(E.swap_acc, z.data(), nrep * sizeof(unsigned long long), hipMemcpyHostToDevice) != hipSuccess;
        if (rc) efail("initial copies failed");
    }
    if (!rc) rc = build_tables(X);
    if (rc) {
        std::string keep = e_err;
        chll_ensemble_destroy(X);
        e_err = keep;
        return -1;
    }
    X->stream = X->own;
    *out = X;
    return 0;
}

int chll_ensemble_destroy(chll_ensemble *X)
{
    if (!X) return 0;
    if (X->stream) (void)hipStreamSynchronize(X->stream);
    drop_graph(X);
    for (RepDesc &D : X->rep) {
        (void)hipFree(D.F.nx); (void)hipFree(D.F.ny); (void)hipFree(D.F.nz); (void)hipFree(D.F.s);
        (void)hipFree(D.counts);
        if (D.flipbits) (void)hipFree(D.flipbits);
        if (D.held) (void)hipFree(D.held);
    }
    EnsDev &E = X->E;
    void *ptrs[] = {X->R, X->blk_rep, X->blk_rep2, X->blk_rep3, X->order_dev, E.kbt, E.inv_kbt, E.d, E.rep_slot,
                    E.slot_rep, (void *)E.lad_start, E.E, E.istep, E.row, E.swap_att, E.swap_acc, E.part, E.spart,
                    E.rows, X->stage, X->stage_s, X->bad};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (X->pinned) (void)hipHostFree(X->pinned);
    for (int q = 0; q < chll_ensemble::kMaxBranches; ++q) {
        if (X->br[q]) (void)hipStreamDestroy(X->br[q]);
        if (X->ev_join[q]) (void)hipEventDestroy(X->ev_join[q]);
    }
    if (X->ev_fork) (void)hipEventDestroy(X->ev_fork);
    if (X->own) (void)hipStreamDestroy(X->own);
    delete X;
    return 0;
}

int chll_ensemble_set_stream(chll_ensemble *X, void *s)
{
    if (!X) return efail("null ensemble");
    X->stream = s ? (hipStream_t)s : X->own;
    drop_graph(X);
    return 0;
}

int chll_ensemble_set_fusion(chll_ensemble *X, int32_t on)
{
    if (!X) return efail("null ensemble");
    if (on < 0 || on > 2) return efail("fusion must be 0 (off), 1 (on) or 2 (on, one site per thread)");
    X->fuse = on;
    return 0;
}

int chll_ensemble_set_chunk(chll_ensemble *X, int64_t bytes)
{
    if (!X) return efail("null ensemble");
    if (bytes < 0) return efail("chunk size must be >= 0 bytes (0: one chunk)");
    ECHECK(hipStreamSynchronize(X->stream));
    X->chunk_bytes = (size_t)bytes;
    build_chunks(X);
    drop_graph(X);
    return 0;
}

int32_t chll_ensemble_chunks(chll_ensemble *X) { return X ? (int32_t)X->chunks.size() : -1; }

int chll_ensemble_set_branches(chll_ensemble *X, int32_t n)
{
    if (!X) return efail("null ensemble");
    if (n < 1 || n > chll_ensemble::kMaxBranches) return efail("branches must be 1 to %d", chll_ensemble::kMaxBranches);
    ECHECK(hipStreamSynchronize(X->stream));
    if (n > 1 && !X->ev_fork) ECHECK(hipEventCreateWithFlags(&X->ev_fork, hipEventDisableTiming));
    for (int q = 0; q < n && n > 1; ++q) {
        if (!X->br[q]) ECHECK(hipStreamCreateWithFlags(&X->br[q], hipStreamNonBlocking));
        if (!X->ev_join[q]) ECHECK(hipEventCreateWithFlags(&X->ev_join[q], hipEventDisableTiming));
    }
    X->branches = n;
    drop_graph(X);
    return 0;
}

int chll_ensemble_set_block(chll_ensemble *X, int32_t bs)
{
    if (!X) return efail("null ensemble");
    if (bs != 64 && bs != 128 && bs != 256 && bs != 512) return efail("block size must be 64, 128, 256 or 512");
    ECHECK(hipStreamSynchronize(X->stream));
    if (bs == X->bs) return 0;
    X->bs = bs;
    return build_tables(X);
}

static int ens_stage(chll_ensemble *X, size_t n)
{
    if (n <= X->stage_n) return 0;
    if (X->stage) (void)hipFree(X->stage);
    if (X->stage_s) (void)hipFree(X->stage_s);
    X->stage = nullptr;
    X->stage_s = nullptr;
    X->stage_n = 0;
    if (ealloc(&X->stage, n) || ealloc(&X->stage_s, n)) return -1;
    X->stage_n = n;
    return 0;
}

int chll_ensemble_upload(chll_ensemble *X, int32_t r, const double *nx, const double *ny, const double *nz,
                         const int32_t *s)
{
    if (!X || r < 0 || r >= X->nrep || !nx || !ny || !nz || !s) return efail("chll_ensemble_upload: bad argument");
    const RepDesc &D = X->rep[r];
    const size_t n3 = 2 * (size_t)D.G.nsub;
    if (ens_stage(X, n3)) return -1;
    const double *src[3] = {nx, ny, nz};
    double *dst[3] = {D.F.nx, D.F.ny, D.F.nz};
    for (int c = 0; c < 3; ++c) {
        ECHECK(hipMemcpyAsync(X->stage, src[c], n3 * sizeof(double), hipMemcpyHostToDevice, X->stream));
        launch_scatter_f64(D.G, X->stage, dst[c], X->stream);
    }
    ECHECK(hipMemsetAsync(X->bad, 0, sizeof(int), X->stream));
    ECHECK(hipMemcpyAsync(X->stage_s, s, n3 * sizeof(int32_t), hipMemcpyHostToDevice, X->stream));
    launch_scatter_spin(D.G, X->stage_s, D.F.s, X->bad, X->stream);
    ECHECK(hipGetLastError());
    ECHECK(hipMemcpyAsync(X->pinned, X->bad, sizeof(int), hipMemcpyDeviceToHost, X->stream));
    ECHECK(hipStreamSynchronize(X->stream));
    int bad;
    memcpy(&bad, X->pinned, sizeof bad);
    if (bad) return efail("spin values must lie in [-127, 127]");
    return 0;
}

int chll_ensemble_init_state(chll_ensemble *X, int32_t kind, uint64_t seed, double p_flip)
{
    if (!X) return efail("null ensemble");
    if (kind < CHLL_INIT_ISOTROPIC || kind > CHLL_INIT_ALIGNED_XZ) return efail("unknown initial state kind %d", kind);
    if (!(p_flip >= 0.0 && p_flip <= 1.0)) return efail("p_flip must lie in [0, 1]");
    for (int r = 0; r < X->nrep; ++r) launch_init_state(X->rep[r].G, X->rep[r].F, kind, seed, (uint32_t)r, p_flip,
                                                          X->stream);
    ECHECK(hipGetLastError());
    ECHECK(hipStreamSynchronize(X->stream));
    return 0;
}

int chll_ensemble_download(chll_ensemble *X, int32_t r, double *nx, double *ny, double *nz, int32_t *s)
{
    if (!X || r < 0 || r >= X->nrep || !nx || !ny || !nz || !s) return efail("chll_ensemble_download: bad argument");
    const RepDesc &D = X->rep[r];
    const size_t n3 = 2 * (size_t)D.G.nsub;
    if (ens_stage(X, n3)) return -1;
    double *dst[3] = {nx, ny, nz};
    const double *src[3] = {D.F.nx, D.F.ny, D.F.nz};
    for (int c = 0; c < 3; ++c) {
        launch_gather_f64(D.G, src[c], X->stage, X->stream);
        ECHECK(hipMemcpyAsync(dst[c], X->stage, n3 * sizeof(double), hipMemcpyDeviceToHost, X->stream));
    }
    launch_gather_spin(D.G, D.F.s, X->stage_s, X->stream);
    ECHECK(hipMemcpyAsync(s, X->stage_s, n3 * sizeof(int32_t), hipMemcpyDeviceToHost, X->stream));
    ECHECK(hipGetLastError());
    ECHECK(hipStreamSynchronize(X->stream));
    return 0;
}

int chll_ensemble_run(chll_ensemble *X, uint64_t seed, uint32_t run_index, int32_t nblocks, int32_t K, int32_t adapt,
                      int32_t use_graph, double *rows, int32_t max_rows)
{
    if (!X) return efail("null ensemble");
    if (nblocks < 0 || K < 1) return efail("need nblocks >= 0 and sweeps per block K >= 1");
    if (max_rows < 0) max_rows = 0;
    hipStream_t st = X->stream;
    EnsDev &E = X->E;
    if (max_rows > X->rows_cap) {
        ECHECK(hipStreamSynchronize(st));
        if (E.rows) (void)hipFree(E.rows);
        E.rows = nullptr;
        X->rows_cap = 0;
        if (ealloc(&E.rows, (size_t)max_rows * X->nrep * kEnsCols)) return -1;
        X->rows_cap = max_rows;
        drop_graph(X);
    }
    // a captured graph bakes in the kernel arguments: rebuild it when they change
    if (X->graph && (X->graph_K != K || X->graph_bs != X->bs || X->graph_adapt != adapt || X->graph_seed != seed ||
                     X->graph_run != run_index || E.max_rows != max_rows || X->graph_fuse != X->fuse))
        drop_graph(X);
    E.seed = seed;
    E.run_index = run_index;
    E.K = (uint32_t)K;
    E.adapt = adapt ? 1 : 0;
    E.max_rows = max_rows;
    const int32_t zero = 0;
    memcpy(X->pinned, &zero, sizeof zero);
    ECHECK(hipMemcpyAsync(E.row, X->pinned, sizeof(int32_t), hipMemcpyHostToDevice, st));
    if (use_graph && nblocks > 0 && !X->gexec) {
        ECHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = launch_block(X, K, st);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return -1;
        }
        if (ce != hipSuccess) return efail("hipStreamEndCapture: %s", hipGetErrorString(ce));
        X->graph = g;
        ECHECK(hipGraphInstantiate(&X->gexec, g, nullptr, nullptr, 0));
        X->graph_K = K;
        X->graph_bs = X->bs;
        X->graph_adapt = adapt;
        X->graph_seed = seed;
        X->graph_run = run_index;
        X->graph_fuse = X->fuse;
    }
    for (int b = 0; b < nblocks; ++b) {
        if (use_graph) ECHECK(hipGraphLaunch(X->gexec, st));
        else if (launch_block(X, K, st)) return -1;
    }
    if (rows && max_rows > 0) {
        const int32_t n = nblocks < max_rows ? nblocks : max_rows;
        if (n > 0)
            ECHECK(hipMemcpyAsync(rows, E.rows, (size_t)n * X->nrep * kEnsCols * sizeof(double),
                                  hipMemcpyDeviceToHost, st));
        ECHECK(hipStreamSynchronize(st));
    }
    return 0;
}

int chll_ensemble_sync(chll_ensemble *X)
{
    if (!X) return efail("null ensemble");
    ECHECK(hipStreamSynchronize(X->stream));
    return 0;
}

int chll_ensemble_state(chll_ensemble *X, int32_t *rep_slot, double *slot_kbt, double *slot_d, int64_t *swap_att,
                        int64_t *swap_acc, uint32_t *istep)
{
    if (!X) return efail("null ensemble");
    hipStream_t st = X->stream;
    const EnsDev &E = X->E;
    const size_t n = X->nrep;
    ECHECK(hipStreamSynchronize(st));
    if (rep_slot) ECHECK(hipMemcpy(rep_slot, E.rep_slot, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (slot_kbt) ECHECK(hipMemcpy(slot_kbt, E.kbt, n * sizeof(double), hipMemcpyDeviceToHost));
    if (slot_d) ECHECK(hipMemcpy(slot_d, E.d, n * sizeof(double), hipMemcpyDeviceToHost));
    if (swap_att) ECHECK(hipMemcpy(swap_att, E.swap_att, n * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (swap_acc) ECHECK(hipMemcpy(swap_acc, E.swap_acc, n * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (istep) ECHECK(hipMemcpy(istep, E.istep, sizeof(uint32_t), hipMemcpyDeviceToHost));
    return 0;
}

int32_t chll_ensemble_blocks(chll_ensemble *X)
{
    if (!X) return efail("null ensemble");
    return (int32_t)X->nblocks;
}

}  // extern "C"
