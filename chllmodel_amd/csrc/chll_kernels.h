// chll_kernels.h -- argument blocks and launchers of the sweep kernels.
#pragma once

#include "chll_device.h"

namespace chll {

// Device fields, colour-split: colour 0 at [0, nsub), colour 1 at [nsub, 2 nsub).
struct Fields {
    double *nx, *ny, *nz;
    int8_t *s;
};

// Per-launch arguments of a half-sweep.
struct PhaseArgs {
    double g, KK, kbt, inv_kbt;  // inv_kbt = RN(1/kbt)
    const double *d;      // device cone step (updated by the finalize kernel)
    const double *u_phi;  // deterministic mode: uniforms indexed by packed site
    const double *u_acc;
    const double *u_spin; // deterministic mode, fused director + chirality launch: chirality uniforms
    uint64_t seed;
    uint32_t run_index, replica, istep;
    uint32_t istep_spin;  // fused chirality colour 0 + next director colour 0: the chirality trial's sweep
    double inv_kbt_spin;  // the same fused launch: RN(1/kbt) of the chirality trial's sweep (kT may ramp)
    double kbt_spin;      // and its kT (reference arithmetic divides by it)
    double screen;        // reference arithmetic: compat_screen's scale (0: every trial rounds both energies)
    // Reference arithmetic with colour 1's chirality trials inside its
    // director launch (held trials, chll_site.h): bit m % 64 of word m / 64
    // is set when colour-1 packed site m flipped in this sweep; null when
    // colour 1's chirality trials have not run yet in this sweep.
    unsigned long long *flipbits;
    // held[0]: count of colour-1 chirality trials the screen left open in
    // this sweep; held[1]: blocks of k_held done; held[kHeldHead + e]: their
    // packed indices (bit 31: decided to flip)
    uint32_t *held;
};

// Held chirality trials: lanes per held trial in k_held (six colour-0
// neighbours and the trial itself), its grid, the queue's header words and
// the flag of a decided flip.
constexpr int kHeldLanes = 8;
constexpr uint32_t kHeldBlocks = 64;
constexpr uint32_t kHeldBlockThreads = 256;
constexpr uint32_t kHeldHead = 2;
constexpr uint32_t kHeldFlip = 0x80000000u;

// compat_screen's scale: 2^-21 bounds the binary32 roundings of two site
// energies (chll_device.h); larger scales only send more trials to the
// rounded energies.
constexpr double kScreenScale = 0x1.0p-21;

constexpr int kRowCols = 16;

// Accept/flip counters: each block adds its count atomically into one of
// kSlots slots of its group (slot = block % kSlots, slots kSlotStride words
// apart so they sit on separate cache lines); the finalize kernel sums and
// re-zeroes them.  Single-lattice groups: director accepts of colours 0 and 1,
// then chirality flips of colours 0 and 1 of even sweeps, then of odd sweeps,
// so that a sweep's colour-0 flips can still be open while the next sweep's
// director trials count (the replica ensemble uses the first four groups as
// director 0, director 1, chirality 0, chirality 1).
constexpr int kSlots = 64;
constexpr int kSlotStride = 32;
constexpr int kGroups = 6;
constexpr int kGroupAcc0 = 0, kGroupAcc1 = 1;
constexpr int kCountWords = kGroups * kSlots * kSlotStride;
__host__ __device__ constexpr int flip_group(int32_t istep, int colour) { return 2 + 2 * (istep & 1) + colour; }
// Cap on measurement blocks (grid-stride), bounding the finalize kernel's sum.
constexpr uint32_t kMeasBlocksMax = 2048;

// The finalize kernel always closes the director trials of sweep istep (sums
// the accept groups, adapts d, keeps the count in pend[istep & 1]); with
// close_prev it also closes the chirality trials of sweep istep - 1 (writes
// its counts_out entry from pend), with close_this those of sweep istep
// (counts_out entry and, with measure, the row).
struct FinArgs {
    uint32_t *counts;        // kGroups x kSlots slots, zeroed after reading
    long long *pend;         // 2 director accept totals awaiting their sweep's flips
    const double *part;      // nblk_meas x 8 measurement partials
    const long long *spart;  // nblk_meas spin partials
    uint32_t nblk_meas;
    uint64_t n3;
    double *d;               // device cone step, in/out
    double *rows;            // rows x kRowCols
    long long *counts_out;   // optional per-sweep (naccept, nflip)
    int32_t istep, measure, row, adapt;
    int32_t close_prev, close_this;
    int64_t sweep_index;     // of sweep istep in counts_out
};

uint32_t phase_blocks(const Geom &G);
uint32_t measure_blocks(const Geom &G);

void launch_scatter_f64(const Geom &G, const double *src, double *dst, hipStream_t s);
void launch_gather_f64(const Geom &G, const double *src, double *dst, hipStream_t s);
void launch_scatter_spin(const Geom &G, const int32_t *src, int8_t *dst, int *bad, hipStream_t s);
void launch_gather_spin(const Geom &G, const int8_t *src, int32_t *dst, hipStream_t s);
// kind 0 = director, 1 = chirality.  order bit 0: XCD row bands (band_block
// with xcd_band blocks per plane group; lattice order when 0); bit 1: two
// sites per thread with 16-byte pair loads where pairs_ok (else one site per
// thread).
void launch_phase(int kind, int arith, int rng, int bc, const Geom &G, const Fields &F, int colour,
                  const PhaseArgs &A, uint32_t *counts, int order, hipStream_t s);
uint32_t xcd_band(const Geom &G);
bool pairs_ok(const Geom &G);
// Director trials of `colour` followed at each site by its chirality trial,
// accepts counted into counts_dir, flips into counts_spin.  Reference
// arithmetic (colour 1, site pairs only: compat_fused_ok): the chirality
// trials the screen leaves open are held (A.held) and every flip is marked
// in A.flipbits.
void launch_phase_fused(int arith, int rng, int bc, const Geom &G, const Fields &F, int colour, const PhaseArgs &A,
                        uint32_t *counts_dir, uint32_t *counts_spin, int order, hipStream_t s);
// Philox: chirality trials of colour 0 for sweep A.istep_spin, each followed
// by the same site's director trial for sweep A.istep (reference arithmetic:
// site pairs only, after the held trials of colour 1 are decided).
void launch_phase_fused_next(int arith, int bc, const Geom &G, const Fields &F, const PhaseArgs &A,
                             uint32_t *counts_dir, uint32_t *counts_spin, int order, hipStream_t s);
// Whether reference-arithmetic sweeps can fuse on this geometry with these
// kernel orders (site pairs for both the director and chirality kernels).
bool compat_fused_ok(const Geom &G, int dir_order, int spin_order);
// Words of the flip bitmap for a geometry (A.flipbits).
size_t flipbit_words(const Geom &G);
// The held chirality trials of colour 1 (A1: sweep t's colour-1 chirality
// arguments; A0: its colour-0 ones): decide each from colour 0's outcomes,
// count the flips into counts_spin, then apply them and clear the queue.
// With fin, the same launch also does that finalize kernel's work (it reads
// no counter the held trials write).
void launch_held(int rng, int bc, const Geom &G, const Fields &F, const PhaseArgs &A0, const PhaseArgs &A1,
                 uint32_t *counts_spin, unsigned long long *total, const FinArgs *fin, hipStream_t s);
// Clears the held-trial queue (a kernel, so that a captured call records no memset).
void launch_held_clear(uint32_t *held, hipStream_t s);
// Whether every field (2 nsub + kFieldPad elements of at most 8 bytes) spans
// less than 4 GiB, so that kernels may address it with 32-bit byte offsets.
bool field_off32(const Geom &G);
// k-marching half-sweeps through LDS (chll_march.hip); tile 0: 64x4, 1: 32x8,
// 2: 16x16 packed sites x rows; returns -1 for an unknown tile.
int launch_phase_march(int kind, int arith, int rng, int bc, int tile, const Geom &G, const Fields &F, int colour,
                       const PhaseArgs &A, uint32_t *counts, uint32_t nslab, hipStream_t s);
uint32_t march_tiles(int tile, const Geom &G);
void launch_measure(int bc, const Geom &G, const Fields &F, double g, double KK, double *part, long long *spart,
                    hipStream_t s);
void launch_finalize(const FinArgs &A, hipStream_t s);
void launch_stream_fill(double *out, uint64_t seed, uint64_t base, uint32_t nsub, hipStream_t s);
// Initial state of every site (k_init_state): kind 0 isotropic directors, 1
// aligned along x, 2 aligned along (1, 0, 1)/sqrt(2); spins -1 with
// probability p_flip.
void launch_init_state(const Geom &G, const Fields &F, int kind, uint64_t seed, uint32_t replica, double p_flip,
                       hipStream_t s);
// One pass of a streaming pattern over `bytes` of each array (k_bandwidth):
// 0 copy a -> b, 1 read a and b and rewrite a, 2 read a, 3 write b.  -1 on a
// bad pattern or size.
int launch_bandwidth(int pattern, void *a, void *b, uint64_t bytes, hipStream_t s);

}  // namespace chll
