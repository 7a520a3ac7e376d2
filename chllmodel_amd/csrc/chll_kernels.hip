// chll_kernels.hip -- HIP kernels (gfx950) of the chiral Lebwohl-Lasher
// Metropolis sweep and their host-side launchers.
//
// One sweep = four dependent half-sweeps over the colour-split lattice
// (director colour 0, director colour 1, chirality colour 0, chirality colour
// 1; chll/chll_mod.f90:70-81), then a one-block finalize kernel that sums the
// per-block accept counters, adapts the cone step d on the device
// (chll/chll_mod.f90:83-90) and, on measurement sweeps, closes the
// observable reductions into one row (chll/chll_mod.f90:91-105).  In FP64
// arithmetic the half-sweeps pair up into fused launches (k_dir_spin /
// k_pair<2>: director + chirality of colour 1; k_spin_dir / k_pair<3>:
// chirality of colour 0 + the next sweep's director of colour 0).  Block order
// (sweep_block): XCD row bands, and consecutive launches alternate direction.
// Nothing in a run of sweeps waits for the host.
#include "chll_kernels.h"
#include "chll_site.h"

namespace chll {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// Layout transforms: Fortran-ordered host fields <-> colour-split SoA.
__global__ __launch_bounds__(kBlock) void k_scatter_f64(Geom G, const double *__restrict__ src,
                                                        double *__restrict__ dst)
{
    const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= G.nsub) return;
    const Site s0 = decode_l(G, 2u * m);
    const uint32_t c0 = (s0.i + s0.j + s0.k) & 1u;  // colour of site l = 2m; site 2m + 1 has the other
    const double a = src[2ull * m], b = src[2ull * m + 1];
    dst[c0 * G.nsub + m] = a;
    dst[(c0 ^ 1u) * G.nsub + m] = b;
}

__global__ __launch_bounds__(kBlock) void k_gather_f64(Geom G, const double *__restrict__ src,
                                                       double *__restrict__ dst)
{
    const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= G.nsub) return;
    const Site s0 = decode_l(G, 2u * m);
    const uint32_t c0 = (s0.i + s0.j + s0.k) & 1u;
    dst[2ull * m] = src[c0 * G.nsub + m];
    dst[2ull * m + 1] = src[(c0 ^ 1u) * G.nsub + m];
}

__global__ __launch_bounds__(kBlock) void k_scatter_spin(Geom G, const int32_t *__restrict__ src,
                                                         int8_t *__restrict__ dst, int *__restrict__ bad)
{
    const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= G.nsub) return;
    const Site s0 = decode_l(G, 2u * m);
    const uint32_t c0 = (s0.i + s0.j + s0.k) & 1u;
    const int32_t a = src[2ull * m], b = src[2ull * m + 1];
    // int8 storage holds the spin exactly when |s| <= 127 (so that -s fits too).
    if (a < -127 || a > 127 || b < -127 || b > 127) atomicOr(bad, 1);
    dst[c0 * G.nsub + m] = (int8_t)a;
    dst[(c0 ^ 1u) * G.nsub + m] = (int8_t)b;
}

__global__ __launch_bounds__(kBlock) void k_gather_spin(Geom G, const int8_t *__restrict__ src,
                                                        int32_t *__restrict__ dst)
{
    const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= G.nsub) return;
    const Site s0 = decode_l(G, 2u * m);
    const uint32_t c0 = (s0.i + s0.j + s0.k) & 1u;
    dst[2ull * m] = src[c0 * G.nsub + m];
    dst[2ull * m + 1] = src[(c0 ^ 1u) * G.nsub + m];
}

// ---------------------------------------------------------------------------
// Director half-sweep over one colour (chll/chll_mod.f90:117-202); block order
// sweep_block (chll_device.h).

template <int ARITH, int RNG, int BC, bool OFF32>
__global__ __launch_bounds__(kBlock) void k_dir(Geom G, Fields F, int colour, PhaseArgs A,
                                                uint32_t *__restrict__ counts, uint32_t band)
{
    const uint32_t blk = sweep_block(band), m = blk * kBlock + threadIdx.x;
    const bool acc = blk != kIdleBlock && m < G.nsub && dir_site<ARITH, RNG, BC, OFF32>(G, F, colour, A, m);
    block_count<kBlock>(acc, counts, blockIdx.x);
}

// Chirality half-sweep over one colour (chll/chll_mod.f90:204-245).
template <int ARITH, int RNG, int BC, bool OFF32>
__global__ __launch_bounds__(kBlock) void k_spin(Geom G, Fields F, int colour, PhaseArgs A,
                                                 uint32_t *__restrict__ counts, uint32_t band)
{
    const uint32_t blk = sweep_block(band), m = blk * kBlock + threadIdx.x;
    const bool acc = blk != kIdleBlock && m < G.nsub && spin_site<ARITH, RNG, BC, OFF32>(G, F, colour, A, m);
    block_count<kBlock>(acc, counts, blockIdx.x);
}

// Director half-sweep over one colour with each site's chirality trial fused
// behind its director trial (FP64 arithmetic; see dir_spin_site).
template <int RNG, int BC, bool OFF32>
__global__ __launch_bounds__(kBlock) void k_dir_spin(Geom G, Fields F, int colour, PhaseArgs A,
                                                     uint32_t *__restrict__ counts_dir,
                                                     uint32_t *__restrict__ counts_spin, uint32_t band)
{
    const uint32_t blk = sweep_block(band), m = blk * kBlock + threadIdx.x;
    bool flip = false;
    const bool acc = blk != kIdleBlock && m < G.nsub && dir_spin_site<RNG, BC, OFF32>(G, F, colour, A, m, flip);
    block_count2<kBlock>(acc, flip, counts_dir, counts_spin, blockIdx.x);
}

// Chirality half-sweep of colour 0 (sweep A.istep_spin) with each site's
// director trial of the next sweep (A.istep) fused behind it (FP64, Philox;
// see spin_dir_site).
template <int BC, bool OFF32>
__global__ __launch_bounds__(kBlock) void k_spin_dir(Geom G, Fields F, PhaseArgs A,
                                                     uint32_t *__restrict__ counts_dir,
                                                     uint32_t *__restrict__ counts_spin, uint32_t band)
{
    const uint32_t blk = sweep_block(band), m = blk * kBlock + threadIdx.x;
    bool flip = false;
    const bool acc = blk != kIdleBlock && m < G.nsub && spin_dir_site<BC, OFF32>(G, F, A, m, flip);
    block_count2<kBlock>(acc, flip, counts_dir, counts_spin, blockIdx.x);
}

// Two sites per thread (load_pair: 16-byte loads of site pairs).  OP 0:
// director trials, 1: chirality trials, 2: director then chirality trials
// (colour 1 in reference arithmetic), 3: chirality trials of sweep
// A.istep_spin then director trials of sweep A.istep (Philox, colour 0),
// 4: chirality trials of colour 0 in reference arithmetic after colour 1's
// (spin_c0_compat).  Counts: OP 0 accepts and OP 1 and 4 flips into c0; OP 2
// and 3 director accepts into c0, flips into c1.  Reference arithmetic, OP 2: each wave writes the
// flip bits of its 128 sites (two 64-bit words, site m at bit m % 64 of word
// m / 64) and appends its held trials to A.held.  Needs an even half
// (ni % 4 == 0) and fields below 4 GiB.
// Waves per SIMD the compiler must fit: the reference-arithmetic colour-0
// fused launch (OP 3) fits 3 with a few spilled registers and runs 9 % faster
// than at the 2 its natural 176-184 VGPRs give; the colour-1 one (OP 2) would
// spill over 200 bytes per lane.
template <int OP, int ARITH>
constexpr int kPairMinWaves = (OP == 3 && ARITH == ARITH_COMPAT) ? 3 : 1;

template <int OP, int ARITH, int RNG, int BC>
__global__ __launch_bounds__(kBlock, (kPairMinWaves<OP, ARITH>)) void k_pair(Geom G, Fields F, int colour, PhaseArgs A,
                                             uint32_t *__restrict__ c0, uint32_t *__restrict__ c1,
                                             uint32_t band)
{
    const uint32_t blk = sweep_block(band), m0 = 2u * (blk * kBlock + threadIdx.x);
    bool a0 = false, a1 = false, f0 = false, f1 = false, h0 = false, h1 = false;
    const double d = OP == 1 ? 0.0 : *A.d;  // the cone step, read once for both sites
    if (blk != kIdleBlock && m0 < G.nsub) {
        SiteIn S0, S1;
        const uint32_t own = load_pair<BC, true>(G, F, colour, m0, S0, S1);
        // both sites' states go back as 16-byte words (store_pair)
        SiteOut O0, O1;
        if constexpr (OP == 0) {
            a0 = dir_apply<ARITH, RNG, true, true>(F, colour, A, d, m0, own, S0, &O0);
            a1 = dir_apply<ARITH, RNG, true, true>(F, colour, A, d, m0 + 1u, own + 1u, S1, &O1);
        } else if constexpr (OP == 1) {
            a0 = spin_apply<ARITH, RNG, true, true>(F, colour, A, m0, own, S0, &O0);
            a1 = spin_apply<ARITH, RNG, true, true>(F, colour, A, m0 + 1u, own + 1u, S1, &O1);
        } else if constexpr (OP == 4) {
            a0 = spin_c0_apply<RNG, BC>(G, A, m0, S0, &O0);
            a1 = spin_c0_apply<RNG, BC>(G, A, m0 + 1u, S1, &O1);
        } else if constexpr (OP == 2) {
            if constexpr (ARITH == ARITH_COMPAT) {
                a0 = dir_spin_held<RNG>(A, d, m0, S0, f0, h0, &O0);
                a1 = dir_spin_held<RNG>(A, d, m0 + 1u, S1, f1, h1, &O1);
            } else {
                a0 = dir_spin_apply<RNG, true, true>(F, colour, A, d, m0, own, S0, f0, &O0);
                a1 = dir_spin_apply<RNG, true, true>(F, colour, A, d, m0 + 1u, own + 1u, S1, f1, &O1);
            }
        } else {
            if constexpr (ARITH == ARITH_COMPAT) {
                a0 = spin_dir_held<BC>(G, A, d, m0, S0, f0, &O0);
                a1 = spin_dir_held<BC>(G, A, d, m0 + 1u, S1, f1, &O1);
            } else {
                a0 = spin_dir_apply<true, true>(F, A, d, m0, own, S0, f0, &O0);
                a1 = spin_dir_apply<true, true>(F, A, d, m0 + 1u, own + 1u, S1, f1, &O1);
            }
        }
        store_pair<OP != 1 && OP != 4, OP != 0>(F, own, O0, O1);
    }
    if constexpr (OP == 2 && ARITH == ARITH_COMPAT)
        held_epilogue(A, G.nsub, blk != kIdleBlock, blk * (kBlock / 64u) + (threadIdx.x >> 6), m0, f0, f1, h0, h1);
    const uint32_t wa = (uint32_t)(__popcll(__ballot(a0)) + __popcll(__ballot(a1)));
    const uint32_t wf = (uint32_t)(__popcll(__ballot(f0)) + __popcll(__ballot(f1)));
    block_sum2<kBlock>(wa, wf, c0, (OP < 2 || OP == 4) ? c0 : c1, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Observables over the whole lattice, one partial per block.  Range r of
// `chunk` packed indices (a multiple of kBlock) goes to the block that XCD
// r / (grid / 8) runs (workgroup b runs on XCD b % 8), in order: each XCD
// walks one contiguous slab of ranges, so the +y rows of a range and the
// rows of the +z plane, a few ranges later in the same slab, come from that
// XCD's L2, and only the slab's last plane reads the next slab's first.  The
// partial of range r is part[r], so the reduction order does not depend on
// the mapping.
__device__ inline uint32_t measure_range()
{
    const uint32_t b = blockIdx.x, n = gridDim.x;
    return (n & 7u) ? b : (b & 7u) * (n >> 3) + (b >> 3);
}

template <int BC>
__global__ __launch_bounds__(kBlock) void k_measure(Geom G, Fields F, double g, double KK,
                                                    double *__restrict__ part, long long *__restrict__ spart,
                                                    uint32_t chunk)
{
    MeasureAcc acc;
    const uint32_t rb = measure_range();
    const uint64_t beg = (uint64_t)rb * chunk;
    const uint64_t end = beg + chunk < G.nsub ? beg + chunk : G.nsub;
    for (uint64_t m = beg + threadIdx.x; m < end; m += kBlock)
        measure_site<BC>(G, F, g, KK, (uint32_t)m, acc);
    block_reduce_measure<kBlock>(acc, part + (size_t)rb * 8, spart + rb);
}

// The same observables marching along k, four sites of one row per thread
// (measure_march_column).  Block (tile t of kBlock packed pairs of a plane,
// slab s of planes [s ks, s ks + ks)).  A row's +y row is read at the step at
// which the lanes of row j + 1 read it as their own row (down) or +z row
// (up), so it comes from L2.  Slabs alternate direction, so the plane two
// adjacent slabs share is read by both at the same time (both start or both
// end there): the slabs of one XCD are contiguous (workgroup b runs on XCD
// b % 8) and its blocks run tile by tile through all of them.  One partial
// per block, in block order.  Needs an even half and
// fields below 4 GiB.
template <int BC>
__global__ __launch_bounds__(kBlock) void k_measure_march(Geom G, Fields F, double g, double KK,
                                                          double *__restrict__ part, long long *__restrict__ spart,
                                                          uint32_t tiles, uint32_t nslab, uint32_t ks)
{
    const uint32_t b = blockIdx.x;
    uint32_t sl, t;
    if ((nslab & 7u) == 0u) {
        // XCD b % 8 takes a contiguous range of slabs, tile by tile: the blocks
        // in flight cover all its slabs of a few tiles, so the two slabs that
        // share a plane read it at about the same time
        const uint32_t n = b >> 3, spx = nslab >> 3;
        t = n / spx;
        sl = (b & 7u) * spx + n % spx;
    } else {
        t = b / nslab;
        sl = b % nslab;
    }
    const uint32_t quads = G.plane >> 1;  // packed pairs per plane and colour
    const uint32_t qd = t * kBlock + threadIdx.x;
    const uint32_t k0 = sl * ks;
    const uint32_t k1 = k0 + ks < G.nk ? k0 + ks : G.nk;
    MeasureAcc acc;
    if (qd < quads && k0 < k1) {
        // lane + 1 holds the next pair of the plane at every step
        const bool nxt = (threadIdx.x & 63u) != 63u && qd + 1u < quads;
        measure_march_column<BC>(G, F, g, KK, qd, k0, k1, (sl & 1u) == 0u, nxt, acc);
    }
    block_reduce_measure<kBlock>(acc, part + (size_t)b * 8, spart + b);
}

// The same observables with four sites of one row per thread (measure_pair);
// chunk is a multiple of 2 kBlock, so lane l + 1 holds the pair after lane
// l's in every pass.  Needs an even half and fields below 4 GiB.
template <int BC>
__global__ __launch_bounds__(kBlock) void k_measure_pair(Geom G, Fields F, double g, double KK,
                                                         double *__restrict__ part, long long *__restrict__ spart,
                                                         uint32_t chunk)
{
    MeasureAcc acc;
    const uint32_t rb = measure_range();
    const uint64_t beg = (uint64_t)rb * chunk;
    const uint32_t end = (uint32_t)(beg + chunk < G.nsub ? beg + chunk : G.nsub);
    const bool last_lane = (threadIdx.x & 63u) == 63u;
    for (uint64_t m0 = beg + 2u * threadIdx.x; m0 < end; m0 += 2u * kBlock)
        measure_pair<BC>(G, F, g, KK, (uint32_t)m0, !last_lane && m0 + 2u < end, acc);
    block_reduce_measure<kBlock>(acc, part + (size_t)rb * 8, spart + rb);
}

// ---------------------------------------------------------------------------
// Sweep epilogue (one block): accept/flip totals, step adaptation, measurement
// row.  Row layout (kRowCols doubles): istep, E (reference etot), eta,
// paccept, faccept, d, E_true/N, eta_exact, Qxx, Qyy, Qzz, Qxy, Qxz, Qyz
// (means of n_a n_b), naccept, nflip.
constexpr int kFinBlock = 256;

// The finalize kernel's work, for one block of kFinBlock threads (k_finalize,
// and the last block of k_held in a fused reference-arithmetic sweep).
__device__ inline void finalize_block(const FinArgs &A)
{
    __shared__ double red[kFinBlock / 64][11];
    const uint32_t t = threadIdx.x;
    unsigned long long ca = 0, cf = 0, cp = 0;
    const int gt = flip_group(A.istep, 0), gp = flip_group(A.istep - 1, 0);
    // Every slot a thread owns is read before any is cleared, so its loads
    // are in flight together (a clear between them would order them).
    constexpr int kPer = (kGroups * kSlots + kFinBlock - 1) / kFinBlock;
    uint32_t *slot[kPer];
    uint32_t cv[kPer];
    int use[kPer];  // 0 none, 1 accepts, 2 this sweep's flips, 3 the previous sweep's
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t q = t + (uint32_t)r * kFinBlock;
        const int g = (int)(q / kSlots);
        use[r] = q >= (uint32_t)(kGroups * kSlots) ? 0
                 : g < 2                                                  ? 1
                 : A.close_this && (g == gt || g == gt + 1)               ? 2
                 : A.close_prev && (g == gp || g == gp + 1)               ? 3
                                                                          : 0;
        slot[r] = use[r] ? A.counts + g * (kSlots * kSlotStride) + (q % kSlots) * kSlotStride : A.counts;
        cv[r] = use[r] ? *slot[r] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        if (!use[r]) continue;
        *slot[r] = 0u;
        if (use[r] == 1) ca += cv[r];
        else if (use[r] == 2) cf += cv[r];
        else cp += cv[r];
    }
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = 0.0;
    long long ss = 0;
    if (A.measure) {
        for (uint32_t b = t; b < A.nblk_meas; b += kFinBlock) {
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] += A.part[(size_t)b * 8 + q];
            ss += A.spart[b];
        }
    }
    double cad = (double)ca, cfd = (double)cf, cpd = (double)cp, ssd = (double)ss;  // exact: < 2^53
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        cad += __shfl_xor(cad, off, 64);
        cfd += __shfl_xor(cfd, off, 64);
        cpd += __shfl_xor(cpd, off, 64);
    }
    if (A.measure) {  // uniform: sweeps without a row skip the observable shuffles
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] += __shfl_xor(v[q], off, 64);
            ssd += __shfl_xor(ssd, off, 64);
        }
    }
    const uint32_t lane = t & 63u, wid = t >> 6;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) red[wid][q] = v[q];
        red[wid][8] = cad;
        red[wid][9] = cfd;
        red[wid][10] = cpd;
    }
    __shared__ double sred[kFinBlock / 64];
    if (lane == 0) sred[wid] = ssd;
    __syncthreads();
    if (t != 0) return;
    double tot[11];
#pragma unroll
    for (int q = 0; q < 11; ++q) {
        double acc = 0.0;
#pragma unroll
        for (int w = 0; w < kFinBlock / 64; ++w) acc += red[w][q];
        tot[q] = acc;
    }
    double stot = 0.0;
#pragma unroll
    for (int w = 0; w < kFinBlock / 64; ++w) stot += sred[w];
    const double n3 = (double)A.n3;
    const double n3f = (double)(float)A.n3;
    // paccept = float(naccept)/float(n3) in binary32 (chll/chll_mod.f90:83);
    // a binary64 quotient of two binary32 values rounds to the binary32 quotient.
    const int32_t nacc = (int32_t)(long long)tot[8];
    const int32_t nflp = (int32_t)(long long)tot[9];
    const double paccept = (double)(float)((double)(float)nacc / n3f);
    double d = *A.d;
    if (A.adapt && A.istep < 10000) {
        if (paccept < 0.4) d = d * 0.99500000476837158203125;
        else if (paccept > 0.6) d = d / 0.99500000476837158203125;
    }
    *A.d = d;
    if (A.close_prev && A.counts_out) {
        A.counts_out[2 * (size_t)(A.sweep_index - 1)] = A.pend[(A.istep - 1) & 1];
        A.counts_out[2 * (size_t)(A.sweep_index - 1) + 1] = (long long)tot[10];
    }
    A.pend[A.istep & 1] = (long long)tot[8];
    if (A.close_this && A.counts_out) {
        A.counts_out[2 * (size_t)A.sweep_index] = (long long)tot[8];
        A.counts_out[2 * (size_t)A.sweep_index + 1] = (long long)tot[9];
    }
    if (A.close_this && A.measure) {
        double *r = A.rows + (size_t)A.row * kRowCols;
        const double faccept = (double)(float)((double)(float)nflp / n3f);
        const long long S = (long long)stot;
        r[0] = (double)A.istep;
        r[1] = (double)(float)(tot[0] / n3f);
        r[2] = (double)(float)((double)(float)S / n3f);
        r[3] = paccept;
        r[4] = faccept;
        r[5] = d;
        r[6] = tot[1] / n3;
        r[7] = (double)S / n3;
        r[8] = tot[2] / n3;
        r[9] = tot[3] / n3;
        r[10] = tot[4] / n3;
        r[11] = tot[5] / n3;
        r[12] = tot[6] / n3;
        r[13] = tot[7] / n3;
        r[14] = tot[8];
        r[15] = tot[9];
    }
}

__global__ __launch_bounds__(kFinBlock) void k_finalize(FinArgs A) { finalize_block(A); }

// Held chirality trials (held_group) of the single lattice.  With fin, one
// more block (the last) runs the finalize kernel's work (finalize_block) for a
// fused sweep's chain: it reads the director accepts of the sweep and the
// flips of the sweep before, none of which the held trials write, so the two
// share a launch.
template <int RNG, int BC>
__global__ __launch_bounds__(kBlock) void k_held(Geom G, Fields F, PhaseArgs A0, PhaseArgs A1,
                                             uint32_t *__restrict__ held, uint32_t *__restrict__ counts_spin,
                                             unsigned long long *__restrict__ total, FinArgs FA, int fin)
{
    const uint32_t nheld_blocks = gridDim.x - (fin ? 1u : 0u);
    if (blockIdx.x >= nheld_blocks) {
        finalize_block(FA);
        return;
    }
    held_group<RNG, BC>(G, F, A0, A1, held, counts_spin, total, blockIdx.x, nheld_blocks);
}


// ---------------------------------------------------------------------------
// Deterministic-mode uniforms for one sweep: the six chunks the reference's
// four half-sweeps consume out of its 10*nsub draws (chll/chll_mod.f90:66-80):
// dir(sl1) phi = [0,n), acc = [n,2n); dir(sl2) phi = [3n,4n), acc = [4n,5n);
// spin(sl1) acc = [7n,8n); spin(sl2) acc = [9n,10n), relative to base.
__global__ __launch_bounds__(kBlock) void k_stream_fill(double *__restrict__ out, uint64_t seed,
                                                        uint64_t base, uint32_t nsub)
{
    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t q = blockIdx.y;
    if (r >= nsub) return;
    const uint64_t off[6] = {0, 1, 3, 4, 7, 9};
    out[(size_t)q * nsub + r] = stream_uniform(seed, base + off[q] * nsub + r);
}

// ---------------------------------------------------------------------------
// Initial states on the device (the reference builds them in Python loops:
// scripts/pdanneal.py:5-16 isotropic, scripts/hero480.py:18-21 aligned xz,
// chll/chll.py:84-98 +-1 spins and the default n = (1, 0, 0)).  Site of
// Fortran index l draws two Philox blocks, key (seed_lo, seed_hi), counters
// (l, replica, kInitTag, 0) and (l, replica, kInitTag, 1):
//   directors (INIT_ISOTROPIC): costh = 2 u1 - 1, u1 = u53(x0, y0);
//     phi = 2 pi u2, u2 = u53(z0, w0); n = (sinth cos phi, sinth sin phi,
//     costh), sinth = sqrt(1 - costh^2), the scripts' formula;
//   spin: -1 when u53(x1, y1) < p_flip, else +1 (chll/chll.py:85-90).
// INIT_ALIGNED_X gives n = (1, 0, 0), INIT_ALIGNED_XZ n = (1, 0, 1)/sqrt(2).
constexpr uint32_t kInitTag = 0x494E4954u;  // "INIT"

__global__ __launch_bounds__(kBlock) void k_init_state(Geom G, Fields F, int kind, uint64_t seed, uint32_t replica,
                                                       double p_flip)
{
    const uint32_t l = blockIdx.x * kBlock + threadIdx.x;
    if (l >= 2u * G.nsub) return;
    const Site st = decode_l(G, l);
    const uint32_t q = ((st.i + st.j + st.k) & 1u) * G.nsub + (l >> 1);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    double nx, ny, nz;
    if (kind == 0) {
        const U4 w = philox4x32_10(U4{l, replica, kInitTag, 0u}, k0, k1);
        const double costh = 2.0 * (u53(w.x, w.y) - 0.5);
        const double sinth = sqrt(1.0 - costh * costh);
        double sphi, cphi;
        sincos(u53(w.z, w.w) * kTwoPi, &sphi, &cphi);
        nx = sinth * cphi; ny = sinth * sphi; nz = costh;
    } else if (kind == 1) {
        nx = 1.0; ny = 0.0; nz = 0.0;
    } else {
        nx = 0.70710678118654752440; ny = 0.0; nz = 0.70710678118654752440;
    }
    const U4 v = philox4x32_10(U4{l, replica, kInitTag, 1u}, k0, k1);
    F.nx[q] = nx; F.ny[q] = ny; F.nz[q] = nz;
    F.s[q] = u53(v.x, v.y) < p_flip ? (int8_t)-1 : (int8_t)1;
}

// ---------------------------------------------------------------------------
// Launchers.
static inline dim3 grid1(uint32_t n) { return dim3((n + kBlock - 1) / kBlock); }

uint32_t phase_blocks(const Geom &G) { return (G.nsub + kBlock - 1) / kBlock; }

// Shape of k_measure_march: tiles of kBlock packed pairs per plane, slabs of
// ks planes (at least 8 where nk allows, more where the blocks would pass
// kMeasBlocksMax).
static void measure_march_shape(const Geom &G, uint32_t &tiles, uint32_t &nslab, uint32_t &ks)
{
    tiles = ((G.plane >> 1) + kBlock - 1) / kBlock;
    const uint32_t cap = kMeasBlocksMax / tiles ? kMeasBlocksMax / tiles : 1u;
    ks = (G.nk + cap - 1) / cap;
    if (ks < 8u) ks = 8u;
    if (ks > G.nk) ks = G.nk;
    nslab = (G.nk + ks - 1) / ks;
}

uint32_t measure_blocks(const Geom &G)
{
    if (pairs_ok(G) && (G.plane >> 1) <= kBlock * kMeasBlocksMax) {
        uint32_t tiles, nslab, ks;
        measure_march_shape(G, tiles, nslab, ks);
        return tiles * nslab;
    }
    const uint32_t b = phase_blocks(G);
    return b < kMeasBlocksMax ? b : kMeasBlocksMax;
}

void launch_scatter_f64(const Geom &G, const double *src, double *dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_scatter_f64, grid1(G.nsub), dim3(kBlock), 0, s, G, src, dst);
}
void launch_gather_f64(const Geom &G, const double *src, double *dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_f64, grid1(G.nsub), dim3(kBlock), 0, s, G, src, dst);
}
void launch_scatter_spin(const Geom &G, const int32_t *src, int8_t *dst, int *bad, hipStream_t s)
{
    hipLaunchKernelGGL(k_scatter_spin, grid1(G.nsub), dim3(kBlock), 0, s, G, src, dst, bad);
}
void launch_gather_spin(const Geom &G, const int8_t *src, int32_t *dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_spin, grid1(G.nsub), dim3(kBlock), 0, s, G, src, dst);
}

// XCD row bands (band_block) for a plane of `units` threads: P = floor(units /
// kBlock) blocks per group (blocks straddle planes when units is no multiple
// of kBlock; group p + 1 then starts less than a block away from plane p + 1,
// so the bands still line up with their z neighbours).  0 (lattice order) when
// a plane holds fewer than 8 blocks or padding its 8 bands to equal size would
// add more than 1/16 to the grid.
static uint32_t plane_band(uint64_t units)
{
    const uint64_t P = units / kBlock, per = (P + 7u) / 8u;
    return P >= 8u && 16u * 8u * per <= 17u * P ? (uint32_t)P : 0u;
}

// Grid of a launch over nblk logical blocks in bands of `band` (bit 31 clear).
static inline dim3 band_grid(uint32_t nblk, uint32_t band)
{
    if (band == 0u) return dim3(nblk);
    const uint32_t groups = (nblk + band - 1u) / band;
    return dim3(groups * 8u * ((band + 7u) / 8u));
}

uint32_t xcd_band(const Geom &G)
{
    // ni even: a plane holds ni nj / 2 packed sites of each colour
    return (G.ni & 1u) ? 0u : plane_band(G.plane);
}

bool field_off32(const Geom &G)
{
    return (2ull * G.nsub + kFieldPad) * sizeof(double) <= 0xffffffffull;
}

bool pairs_ok(const Geom &G) { return (G.ni & 3u) == 0u && field_off32(G); }

// Pair kernels: nsub / 2 threads, a plane of plane / 2 pairs.
static uint32_t pair_band(const Geom &G, int order) { return (order & 1) ? plane_band(G.plane / 2u) : 0u; }
static inline dim3 grid_pairs(const Geom &G, uint32_t band)
{
    return band_grid(grid1(G.nsub / 2u).x, band & 0x7fffffffu);
}

template <int OP, int ARITH, int RNG, int BC>
static void launch_pair(const Geom &G, const Fields &F, int colour, const PhaseArgs &A, uint32_t *c0, uint32_t *c1,
                        uint32_t band, hipStream_t s)
{
    hipLaunchKernelGGL((k_pair<OP, ARITH, RNG, BC>), grid_pairs(G, band), dim3(kBlock), 0, s, G, F, colour, A, c0, c1, band);
}
static uint32_t site_band(const Geom &G, int order) { return (order & 1) ? xcd_band(G) : 0u; }

// order bit 0: XCD row bands, bit 1: site pairs, bit 2: reverse block order.
template <int ARITH, int RNG, int BC>
static void launch_phase_t(int kind, const Geom &G, const Fields &F, int colour, const PhaseArgs &A,
                           uint32_t *counts, int order, hipStream_t s)
{
    const dim3 b(kBlock);
    const uint32_t rev = (order & 4) ? 0x80000000u : 0u;
    if ((order & 2) && pairs_ok(G)) {
        const uint32_t band = pair_band(G, order) | rev;
        if (kind == 0) launch_pair<0, ARITH, RNG, BC>(G, F, colour, A, counts, counts, band, s);
        else if (ARITH == ARITH_COMPAT && colour == 0 && A.flipbits)
            launch_pair<4, ARITH_COMPAT, RNG, BC>(G, F, colour, A, counts, counts, band, s);
        else launch_pair<1, ARITH, RNG, BC>(G, F, colour, A, counts, counts, band, s);
        return;
    }
    const uint32_t band = site_band(G, order) | rev;
    const dim3 g = band_grid(grid1(G.nsub).x, band & 0x7fffffffu);
    if (field_off32(G)) {
        if (kind == 0) hipLaunchKernelGGL((k_dir<ARITH, RNG, BC, true>), g, b, 0, s, G, F, colour, A, counts, band);
        else hipLaunchKernelGGL((k_spin<ARITH, RNG, BC, true>), g, b, 0, s, G, F, colour, A, counts, band);
    } else {
        if (kind == 0) hipLaunchKernelGGL((k_dir<ARITH, RNG, BC, false>), g, b, 0, s, G, F, colour, A, counts, band);
        else hipLaunchKernelGGL((k_spin<ARITH, RNG, BC, false>), g, b, 0, s, G, F, colour, A, counts, band);
    }
}

template <int RNG, int BC>
static void launch_fused_t(int arith, const Geom &G, const Fields &F, int colour, const PhaseArgs &A, uint32_t *cd,
                           uint32_t *cs, int order, hipStream_t s)
{
    const dim3 b(kBlock);
    if (arith == ARITH_COMPAT) {
        // compat_fused_ok: site pairs, colour 1
        launch_pair<2, ARITH_COMPAT, RNG, BC>(G, F, colour, A, cd, cs, pair_band(G, order), s);
        return;
    }
    if ((order & 2) && pairs_ok(G)) {
        launch_pair<2, ARITH_FP64, RNG, BC>(G, F, colour, A, cd, cs, pair_band(G, order), s);
        return;
    }
    const uint32_t band = site_band(G, order);
    const dim3 g = band_grid(grid1(G.nsub).x, band);
    if (field_off32(G)) hipLaunchKernelGGL((k_dir_spin<RNG, BC, true>), g, b, 0, s, G, F, colour, A, cd, cs, band);
    else hipLaunchKernelGGL((k_dir_spin<RNG, BC, false>), g, b, 0, s, G, F, colour, A, cd, cs, band);
}

void launch_phase_fused(int arith, int rng, int bc, const Geom &G, const Fields &F, int colour, const PhaseArgs &A,
                        uint32_t *counts_dir, uint32_t *counts_spin, int order, hipStream_t s)
{
    if (rng == RNG_PHILOX) {
        if (bc == BC_OPEN)
            launch_fused_t<RNG_PHILOX, BC_OPEN>(arith, G, F, colour, A, counts_dir, counts_spin, order, s);
        else launch_fused_t<RNG_PHILOX, BC_PERIODIC>(arith, G, F, colour, A, counts_dir, counts_spin, order, s);
    } else {
        if (bc == BC_OPEN)
            launch_fused_t<RNG_STREAM, BC_OPEN>(arith, G, F, colour, A, counts_dir, counts_spin, order, s);
        else launch_fused_t<RNG_STREAM, BC_PERIODIC>(arith, G, F, colour, A, counts_dir, counts_spin, order, s);
    }
}

void launch_phase_fused_next(int arith, int bc, const Geom &G, const Fields &F, const PhaseArgs &A,
                             uint32_t *counts_dir, uint32_t *counts_spin, int order, hipStream_t s)
{
    const dim3 b(kBlock);
    if (arith == ARITH_COMPAT) {
        // compat_fused_ok: site pairs; reversed like the FP64 launch
        const uint32_t band = pair_band(G, order) | ((order & 4) ? 0x80000000u : 0u);
        if (bc == BC_OPEN)
            launch_pair<3, ARITH_COMPAT, RNG_PHILOX, BC_OPEN>(G, F, 0, A, counts_dir, counts_spin, band, s);
        else
            launch_pair<3, ARITH_COMPAT, RNG_PHILOX, BC_PERIODIC>(G, F, 0, A, counts_dir, counts_spin, band, s);
        return;
    }
    if ((order & 2) && pairs_ok(G)) {
        // order bit 2: this launch walks its blocks in reverse, so it starts on
        // the planes the colour-1 launch before it finished with
        const uint32_t band = pair_band(G, order) | ((order & 4) ? 0x80000000u : 0u);
        if (bc == BC_OPEN)
            launch_pair<3, ARITH_FP64, RNG_PHILOX, BC_OPEN>(G, F, 0, A, counts_dir, counts_spin, band, s);
        else
            launch_pair<3, ARITH_FP64, RNG_PHILOX, BC_PERIODIC>(G, F, 0, A, counts_dir, counts_spin, band, s);
        return;
    }
    const uint32_t band = site_band(G, order) | ((order & 4) ? 0x80000000u : 0u);
    const dim3 g = band_grid(grid1(G.nsub).x, band & 0x7fffffffu);
    const bool o32 = field_off32(G);
    if (bc == BC_OPEN) {
        if (o32) hipLaunchKernelGGL((k_spin_dir<BC_OPEN, true>), g, b, 0, s, G, F, A, counts_dir, counts_spin, band);
        else hipLaunchKernelGGL((k_spin_dir<BC_OPEN, false>), g, b, 0, s, G, F, A, counts_dir, counts_spin, band);
    } else {
        if (o32)
            hipLaunchKernelGGL((k_spin_dir<BC_PERIODIC, true>), g, b, 0, s, G, F, A, counts_dir, counts_spin, band);
        else
            hipLaunchKernelGGL((k_spin_dir<BC_PERIODIC, false>), g, b, 0, s, G, F, A, counts_dir, counts_spin, band);
    }
}

bool compat_fused_ok(const Geom &G, int dir_order, int spin_order)
{
    return pairs_ok(G) && (dir_order & 2) && (spin_order & 2);
}

size_t flipbit_words(const Geom &G)
{
    // two words per wave of the pair launch's logical blocks
    return (size_t)grid1(G.nsub / 2u).x * (kBlock / 64u) * 2u;
}

// Clears the held-trial queue's header (count and blocks done): a call whose
// launches stopped between a colour-1 launch and its k_held leaves them set.
__global__ void k_held_clear(uint32_t *__restrict__ held)
{
    if (threadIdx.x < kHeldHead) held[threadIdx.x] = 0u;
}

void launch_held_clear(uint32_t *held, hipStream_t s)
{
    hipLaunchKernelGGL(k_held_clear, dim3(1), dim3(64), 0, s, held);
}

template <int RNG, int BC>
static void launch_held_t(const Geom &G, const Fields &F, const PhaseArgs &A0, const PhaseArgs &A1,
                          uint32_t *counts_spin, unsigned long long *total, const FinArgs *fin, hipStream_t s)
{
    const FinArgs FA = fin ? *fin : FinArgs{};
    hipLaunchKernelGGL((k_held<RNG, BC>), dim3(kHeldBlocks + (fin ? 1u : 0u)), dim3(kBlock), 0, s, G, F, A0, A1,
                       A1.held, counts_spin, total, FA, fin ? 1 : 0);
}

void launch_held(int rng, int bc, const Geom &G, const Fields &F, const PhaseArgs &A0, const PhaseArgs &A1,
                 uint32_t *counts_spin, unsigned long long *total, const FinArgs *fin, hipStream_t s)
{
    if (rng == RNG_PHILOX) {
        if (bc == BC_OPEN) launch_held_t<RNG_PHILOX, BC_OPEN>(G, F, A0, A1, counts_spin, total, fin, s);
        else launch_held_t<RNG_PHILOX, BC_PERIODIC>(G, F, A0, A1, counts_spin, total, fin, s);
    } else {
        if (bc == BC_OPEN) launch_held_t<RNG_STREAM, BC_OPEN>(G, F, A0, A1, counts_spin, total, fin, s);
        else launch_held_t<RNG_STREAM, BC_PERIODIC>(G, F, A0, A1, counts_spin, total, fin, s);
    }
}

void launch_phase(int kind, int arith, int rng, int bc, const Geom &G, const Fields &F, int colour,
                  const PhaseArgs &A, uint32_t *counts, int order, hipStream_t s)
{
    const bool philox = (rng == RNG_PHILOX);
    if (arith == ARITH_COMPAT) {
        if (philox) {
            if (bc == BC_OPEN) launch_phase_t<ARITH_COMPAT, RNG_PHILOX, BC_OPEN>(kind, G, F, colour, A, counts, order, s);
            else launch_phase_t<ARITH_COMPAT, RNG_PHILOX, BC_PERIODIC>(kind, G, F, colour, A, counts, order, s);
        } else {
            if (bc == BC_OPEN) launch_phase_t<ARITH_COMPAT, RNG_STREAM, BC_OPEN>(kind, G, F, colour, A, counts, order, s);
            else launch_phase_t<ARITH_COMPAT, RNG_STREAM, BC_PERIODIC>(kind, G, F, colour, A, counts, order, s);
        }
    } else {
        if (philox) {
            if (bc == BC_OPEN) launch_phase_t<ARITH_FP64, RNG_PHILOX, BC_OPEN>(kind, G, F, colour, A, counts, order, s);
            else launch_phase_t<ARITH_FP64, RNG_PHILOX, BC_PERIODIC>(kind, G, F, colour, A, counts, order, s);
        } else {
            if (bc == BC_OPEN) launch_phase_t<ARITH_FP64, RNG_STREAM, BC_OPEN>(kind, G, F, colour, A, counts, order, s);
            else launch_phase_t<ARITH_FP64, RNG_STREAM, BC_PERIODIC>(kind, G, F, colour, A, counts, order, s);
        }
    }
}

void launch_measure(int bc, const Geom &G, const Fields &F, double g, double KK, double *part, long long *spart,
                    hipStream_t s)
{
    // measure_blocks(G) partials (the finalize kernel reads that many), each
    // block a contiguous range; trailing blocks may get an empty one
    const uint32_t nb = measure_blocks(G);
    if (pairs_ok(G) && (G.plane >> 1) <= kBlock * kMeasBlocksMax) {
        uint32_t tiles, nslab, ks;
        measure_march_shape(G, tiles, nslab, ks);
        if (bc == BC_OPEN)
            hipLaunchKernelGGL(k_measure_march<BC_OPEN>, dim3(nb), dim3(kBlock), 0, s, G, F, g, KK, part, spart, tiles,
                               nslab, ks);
        else
            hipLaunchKernelGGL(k_measure_march<BC_PERIODIC>, dim3(nb), dim3(kBlock), 0, s, G, F, g, KK, part, spart,
                               tiles, nslab, ks);
        return;
    }
    uint32_t chunk = (G.nsub + nb - 1) / nb;
    if (pairs_ok(G)) {
        chunk = (chunk + 2 * kBlock - 1) / (2 * kBlock) * (2 * kBlock);
        if (bc == BC_OPEN)
            hipLaunchKernelGGL(k_measure_pair<BC_OPEN>, dim3(nb), dim3(kBlock), 0, s, G, F, g, KK, part, spart, chunk);
        else
            hipLaunchKernelGGL(k_measure_pair<BC_PERIODIC>, dim3(nb), dim3(kBlock), 0, s, G, F, g, KK, part, spart,
                               chunk);
        return;
    }
    chunk = (chunk + kBlock - 1) / kBlock * kBlock;
    if (bc == BC_OPEN)
        hipLaunchKernelGGL(k_measure<BC_OPEN>, dim3(nb), dim3(kBlock), 0, s, G, F, g, KK, part, spart, chunk);
    else
        hipLaunchKernelGGL(k_measure<BC_PERIODIC>, dim3(nb), dim3(kBlock), 0, s, G, F, g, KK, part, spart, chunk);
}

void launch_finalize(const FinArgs &A, hipStream_t s)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, s, A);
}

void launch_init_state(const Geom &G, const Fields &F, int kind, uint64_t seed, uint32_t replica, double p_flip,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_init_state, grid1(2u * G.nsub), dim3(kBlock), 0, s, G, F, kind, seed, replica, p_flip);
}

void launch_stream_fill(double *out, uint64_t seed, uint64_t base, uint32_t nsub, hipStream_t s)
{
    dim3 grid((nsub + kBlock - 1) / kBlock, 6);
    hipLaunchKernelGGL(k_stream_fill, grid, dim3(kBlock), 0, s, out, seed, base, nsub);
}


// ---------------------------------------------------------------------------
// Streaming ceilings of this GPU's memory for the mixes the sweep kernels
// move (chll_bandwidth_probe).  Every thread moves four 16-byte words at
// i = 1024 b + 256 u + t, u = 0..3, all loads issued before the first store;
// one launch covers the arrays once (n16 words, tail guarded).
namespace {
typedef double d2v __attribute__((ext_vector_type(2)));
constexpr int kBwUnroll = 4;

template <int PAT>
__global__ __launch_bounds__(kBlock) void k_bandwidth(d2v *__restrict__ a, d2v *__restrict__ b, uint64_t n16)
{
    const uint64_t base = (uint64_t)blockIdx.x * (kBlock * kBwUnroll) + threadIdx.x;
    d2v va[kBwUnroll], vb[kBwUnroll];
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < kBwUnroll; ++u) {
        const uint64_t i = base + (uint64_t)u * kBlock;
        const bool in = i < n16;
        if constexpr (PAT != 3) va[u] = in ? a[i] : d2v{0.0, 0.0};
        if constexpr (PAT == 1) vb[u] = in ? b[i] : d2v{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < kBwUnroll; ++u) {
        const uint64_t i = base + (uint64_t)u * kBlock;
        if (i >= n16) continue;
        if constexpr (PAT == 0) b[i] = va[u];                  // copy a -> b
        if constexpr (PAT == 1) a[i] = va[u] * 0.5 + vb[u];    // read a, b; rewrite a
        if constexpr (PAT == 2) acc += va[u].x + va[u].y;      // read a
        if constexpr (PAT == 3) b[i] = d2v{1.0, (double)u};    // write b
    }
    // the read-only pattern stores only if the sum hits a value zero-filled
    // memory never produces, which keeps its loads alive
    if constexpr (PAT == 2)
        if (acc == 1.25) b[blockIdx.x] = d2v{acc, acc};
}
}  // namespace

int launch_bandwidth(int pattern, void *a, void *b, uint64_t bytes, hipStream_t s)
{
    const uint64_t n16 = bytes / 16u;
    const uint64_t per = (uint64_t)kBlock * kBwUnroll;
    const uint64_t grid = (n16 + per - 1) / per;
    if (grid == 0 || grid > 0x7fffffffull) return -1;
    d2v *pa = (d2v *)a, *pb = (d2v *)b;
    switch (pattern) {
    case 0: k_bandwidth<0><<<(uint32_t)grid, kBlock, 0, s>>>(pa, pb, n16); break;
    case 1: k_bandwidth<1><<<(uint32_t)grid, kBlock, 0, s>>>(pa, pb, n16); break;
    case 2: k_bandwidth<2><<<(uint32_t)grid, kBlock, 0, s>>>(pa, pb, n16); break;
    case 3: k_bandwidth<3><<<(uint32_t)grid, kBlock, 0, s>>>(pa, pb, n16); break;
    default: return -1;
    }
    return 0;
}

}  // namespace chll
